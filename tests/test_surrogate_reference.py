"""Pins of the CPU reference of the surrogate network (tests/mlp_reference.py),
which the GPU training and inference kernels are checked against."""
import numpy as np

from tests import mlp_reference as ref


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = ref.init_params(1)
    x = rng.uniform(-1, 1, size=(16, 7))
    t = rng.normal(size=(16, 7))
    _, g = ref.loss_and_grad(p, x, t, rounded=False)
    # central differences on a sample of parameters of every layer
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in ref.SHAPES])
    for (name, _), a, b in zip(ref.SHAPES, offs[:-1], offs[1:]):
        for k in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            lp, _ = ref.loss_and_grad(pp, x, t, rounded=False)
            lm, _ = ref.loss_and_grad(pm, x, t, rounded=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (name, k, fd, g[k])


def test_rounded_gradient_is_close_to_exact():
    """BF16 rounding at the GPU's rounding points changes the gradient by
    about the BF16 precision, not more."""
    rng = np.random.default_rng(5)
    p = ref.init_params(2)
    x = rng.uniform(-1, 1, size=(256, 7)).astype(np.float32)
    t = rng.normal(size=(256, 7))
    _, ge = ref.loss_and_grad(p, x, t, rounded=False)
    _, gr = ref.loss_and_grad(p, x, t, rounded=True)
    assert np.linalg.norm(gr - ge) / np.linalg.norm(ge) < 3e-2


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], dtype=np.float32)
    r = ref.bf16(v)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]


def test_neohookean_green_matches_oracle_point_stress():
    """The Green-strain form of the neo-Hookean law (used for the training
    set) against the oracle's neo-Hookean point stress on F = U = sqrt(C)."""
    from oracle import oracle as orc
    rng = np.random.default_rng(9)
    E_young, nu = 1e9, 0.3
    lam = nu * E_young / ((1 + nu) * (1 - 2 * nu))
    mu = E_young / (2 * (1 + nu))
    Ev = rng.uniform(-0.05, 0.05, size=(20, 6))
    S, W = ref.neohookean_green(Ev, lam, mu)
    idx = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for k in range(20):
        C = np.eye(3)
        for m, (i, j) in enumerate(idx):
            C[i, j] += 2 * Ev[k, m]
            if i != j:
                C[j, i] += 2 * Ev[k, m]
        w, V = np.linalg.eigh(C)
        U = V @ np.diag(np.sqrt(w)) @ V.T
        So, _ = orc.point_stress(1, 3, E_young, nu, U - np.eye(3))
        Wo = orc.point_energy(1, 3, E_young, nu, U - np.eye(3))
        Sv = np.array([So[i, j] for i, j in idx])
        assert np.abs(Sv - S[k]).max() <= 1e-9 * np.abs(So).max()
        assert abs(Wo - W[k]) <= 1e-9 * abs(Wo)
    # the linear part is the small-strain limit
    Sl, Wl = ref.linear_green(Ev * 1e-4, lam, mu)
    Sn, Wn = ref.neohookean_green(Ev * 1e-4, lam, mu)
    assert np.abs(Sl - Sn).max() <= 1e-3 * np.abs(Sn).max()


def test_activation_shape():
    """The smoothstep swish: a(0) = 0, a'(0) = 1/2 as SiLU, 0 for z <= -1/(2k),
    z for z >= 1/(2k), and continuously differentiable: its derivative agrees
    with central differences everywhere, also across both knots."""
    k = ref.ACT_K
    assert float(np.float16(k)) == k and abs(k - 1.0 / 6.0) < 1e-4
    knot = 0.5 / k
    z = np.concatenate([np.linspace(-10, 10, 20001), [-knot, knot]])
    assert ref.act(np.array([0.0]))[0] == 0.0
    assert abs(ref.act_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    assert (ref.act(z[z <= -knot]) == 0.0).all()
    assert np.allclose(ref.act(z[z >= knot]), z[z >= knot], rtol=0, atol=1e-12)
    def fd_error(f, df):
        h = 1e-6
        return np.abs((f(z + h) - f(z - h)) / (2 * h) - df(z)).max()

    assert fd_error(ref.act, ref.act_grad) < 1e-6
    # the derivative is continuous: no jump across the knots
    for zk in (-knot, knot):
        assert abs(ref.act_grad(np.array([zk - 1e-9]))[0] - ref.act_grad(np.array([zk + 1e-9]))[0]) < 1e-8
    # planted slips fail: a derivative without the z S'(z) term, and a
    # quadratic S (t^2) whose derivative jumps at the upper knot
    def grad_no_dS(zz):
        t = np.clip(k * zz + 0.5, 0.0, 1.0)
        return t * t * (3.0 - 2.0 * t)
    assert fd_error(ref.act, grad_no_dS) > 0.1
    quad = lambda zz: zz * np.clip(k * zz + 0.5, 0.0, 1.0) ** 2
    above = (quad(knot + 2e-6) - quad(knot + 1e-6)) / 1e-6
    below = (quad(knot - 1e-6) - quad(knot - 2e-6)) / 1e-6
    assert abs(above - below) > 0.5


def test_act_f16_against_act():
    """The FP16 evaluation gives a(z) within 2.5e-3 of max(1, |a|) (about 2.5
    FP16 steps; the rounding of z to FP16 alone moves a by up to
    max|a'| x 2^-11 |z|)."""
    z = np.linspace(-60.0, 60.0, 1200001)
    a = ref.act(z)
    g = ref.act_f16(z)
    rel = np.abs(g - a) / np.maximum(1.0, np.abs(a))
    assert rel.max() < 2.5e-3
    assert (g[z <= -0.5 / ref.ACT_K - 0.1] == 0.0).all()


def test_f16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, -(1.0 + 3 * 2 ** -11), 1.0 + 2 ** -12])
    # ties to even
    assert list(ref.f16(v)) == [1.0, 1.0, 1.0 + 2 ** -9, -(1.0 + 2 ** -9), 1.0]


def test_inference_form_is_the_network():
    """The FP16 inference form equals the exact FP64 network to FP16
    precision, closer than the BF16 training form."""
    rng = np.random.default_rng(8)
    p = ref.init_params(4).astype(np.float32).astype(np.float64)
    x = rng.uniform(-1, 1, size=(512, 7)).astype(np.float32)
    ye, _ = ref.forward(p, x, rounded=False)
    yi = ref.forward_inference(p, x)
    yb, _ = ref.forward(p, x, rounded=True)
    ei = np.linalg.norm(yi - ye) / np.linalg.norm(ye)
    eb = np.linalg.norm(yb - ye) / np.linalg.norm(ye)
    assert ei < 2e-3
    assert ei < eb  # closer than the BF16 training form



def test_act_bf16_against_act():
    """The BF16 inference kernel's activation (FP32 operations, one rounding
    to BF16) gives a(z) within half a BF16 step (2^-8) of max(1, |a|), and is
    the BF16 rounding of the exact a(z) except where the FP32 evaluation
    lands on the other side of a rounding boundary (a few points in 10^6)."""
    z = np.linspace(-60.0, 60.0, 1200001)
    a = ref.act(z)
    g = ref.act_bf16(z)
    rel = np.abs(g - a) / np.maximum(1.0, np.abs(a))
    assert rel.max() <= 2.0 ** -8 * 1.0001
    assert np.mean(g != ref.bf16(a).astype(np.float64)) < 1e-5
    # a planted slip in the FP32 steps (3 - t instead of 3 - 2t) is seen
    t = np.clip(ref.f32(ref.f32(z) * ref.ACT_K + 0.5), 0.0, 1.0)
    slip = ref.bf16(ref.f32(z * ref.f32(ref.f32(t * t) * (3.0 - t)))).astype(np.float64)
    assert (np.abs(slip - a) / np.maximum(1.0, np.abs(a))).max() > 0.1


def test_bf16_inference_form_is_the_training_form():
    """The BF16 inference form is the BF16 training forward (weights, inputs
    and activations rounded at the same places) up to the activation's FP32
    evaluation, and the FP16 form is closer to the exact network than both."""
    rng = np.random.default_rng(8)
    p = ref.init_params(4).astype(np.float32).astype(np.float64)
    x = rng.uniform(-1, 1, size=(512, 7)).astype(np.float32)
    ye, _ = ref.forward(p, x, rounded=False)
    yb, _ = ref.forward(p, x, rounded=True)
    yi = ref.forward_inference(p, x, "bf16")
    yf = ref.forward_inference(p, x, "fp16")
    assert np.linalg.norm(yi - yb) / np.linalg.norm(yb) < 3e-4
    ei = np.linalg.norm(yi - ye) / np.linalg.norm(ye)
    ef = np.linalg.norm(yf - ye) / np.linalg.norm(ye)
    assert ei < 1e-2
    assert ef < 0.5 * ei
