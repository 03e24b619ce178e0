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


def test_silu_polynomial_error_bound():
    """Q stands for -2 / (1 + X) within |y| X |dQ| <= 1.14e-4 over the whole
    range (the stated bound of the inference kernel's SiLU); a sign or a 1 %
    slip in one coefficient breaks it (by 2.5x or more)."""
    y = np.linspace(0.0, 80.0, 800001)
    assert ref.silu_poly_error(y).max() <= 1.14e-4
    X = 2.0 ** -y
    for k in range(5):
        for bad in (-ref.SILU_Q[k], ref.SILU_Q[k] * 1.01):
            c = list(ref.SILU_Q)
            c[k] = bad
            Q = sum(ck * X ** j for j, ck in enumerate(c))
            assert (y * X * np.abs(Q + 2.0 / (1.0 + X))).max() > 2.9e-4
    # every coefficient is an FP16 value (the kernel holds them as such)
    assert all(float(np.float16(c)) == c for c in ref.SILU_Q)


def test_silu_f16_against_silu():
    """The FP16 evaluation gives g = 2 log2(e) SiLU(z) to FP16 precision:
    within 1.5e-3 of max(1, |g|) everywhere (4e-4 rms for |z| <= 8), and
    below 5e-5 in absolute value for z < -8, where SiLU itself is tiny."""
    z = np.linspace(-60.0, 60.0, 1200001)
    y = z / np.log(2.0)
    exact = 2.0 * y / (1.0 + np.exp(-z))
    g = ref.silu_f16(y)
    rel = np.abs(g - exact) / np.maximum(1.0, np.abs(exact))
    near = np.abs(z) <= 8.0
    assert rel.max() < 1.5e-3
    assert np.sqrt((rel[near] ** 2).mean()) < 4e-4
    assert np.abs(g - exact)[z < -8.0].max() < 5e-5
    assert (ref.silu_f16(np.array([-30.0, -100.0, -1e4])) == 0.0).all()


def test_f16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, -(1.0 + 3 * 2 ** -11), 1.0 + 2 ** -12])
    # ties to even
    assert list(ref.f16(v)) == [1.0, 1.0, 1.0 + 2 ** -9, -(1.0 + 2 ** -9), 1.0]


def test_inference_form_is_the_network():
    """The scaled FP16 inference form equals the exact FP64 network to FP16
    precision: the scale factors cancel (log2 e into layer 0, 1/2 per hidden
    layer, ln 2 / 2 into the output layer)."""
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
    # a wrong scale (no log2 e on the biases) is caught
    P = ref.unpack(p)
    P["b1"] = P["b1"] / float(ref.SCALE_IN)
    yw = ref.forward_inference(ref.pack(P), x)
    assert np.linalg.norm(yw - ye) / np.linalg.norm(ye) > 1e-2
