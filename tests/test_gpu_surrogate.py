"""The constitutive surrogate on the GPU: MFMA inference and training kernels
against the NumPy reference (tests/mlp_reference.py), the training set
against the neo-Hookean law, what training gains, and the surrogate element
step against the plain neohookean_damage step."""
import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from akantu_amd import rng as akrng
from tests import mlp_reference as ref

pytestmark = pytest.mark.gpu

PLATE = dict(E=1e9, nu=0.3, rho=1200.0)
SD = 5e6


def _lame(E, nu):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def _plate_model(nx=12, ny=12, nz=3, seed=4, eps=0.05):
    m = akm.notched_plate_tet4(nx, ny, nz)
    yd = akrng.weibull_thresholds(m.nb_elements, 1e6, 5.0, seed=seed)
    model = akx.SolidMechanicsModel(m)
    model.add_material(kind=akx.NEOHOOKEAN_DAMAGE, Sd=SD, Yd=yd, **PLATE)
    model.init_full()
    model.surrogate_setup(0, eps)
    return model


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_inference_kernel_matches_reference():
    """Chained 32x32x16 FP16 MFMA layers with the permuted weight image give
    the reference's inference form of the network (checks the operand /
    result lane layout, the bias operand and the packed FP16 activation), and
    the exact FP64 network to FP16 precision."""
    model = _plate_model()
    p = ref.init_params(11).astype(np.float32)
    model.surrogate_params(p)
    assert np.array_equal(model.surrogate_params(), p)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, size=(1000, 7)).astype(np.float32)  # not a multiple of 32
    x[:, 6] = np.abs(x[:, 6])  # damage input in [0, 1]
    y = model.surrogate_infer(x)
    yi = ref.forward_inference(p.astype(np.float64), x)
    # FP32 accumulation against float64: an FP16 rounding of an activation
    # may go the other way now and then
    assert _rel(y, yi) < 5e-4
    assert np.abs(y - yi).max() < 3e-3 * np.abs(yi).max()
    ye, _ = ref.forward(p.astype(np.float64), x, rounded=False)
    assert _rel(y, ye) < 3e-3


def test_bf16_inference_kernel_matches_reference():
    """The inference kernel in BF16 (weights and activations; the training
    format) against the reference's BF16 inference form and the exact
    network; switching back to FP16 rebuilds the FP16 image from the same
    master weights (the same outputs bit for bit)."""
    model = _plate_model()
    p = ref.init_params(11).astype(np.float32)
    model.surrogate_params(p)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, size=(1000, 7)).astype(np.float32)
    x[:, 6] = np.abs(x[:, 6])
    assert model.surrogate_format() == "fp16"
    y16 = model.surrogate_infer(x)
    assert model.surrogate_format("bf16") == "bf16"
    y = model.surrogate_infer(x)
    yi = ref.forward_inference(p.astype(np.float64), x, "bf16")
    assert _rel(y, yi) < 5e-4
    assert np.abs(y - yi).max() < 5e-3 * np.abs(yi).max()
    ye, _ = ref.forward(p.astype(np.float64), x, rounded=False)
    assert _rel(y, ye) < 1e-2
    assert _rel(y16, ye) < _rel(y, ye)
    model.surrogate_format("fp16")
    assert np.array_equal(model.surrogate_infer(x), y16)
    # an unknown format is refused and changes nothing
    assert akx.lib.akx_surrogate_format(model.ctx.h, 7) == -1
    assert model.surrogate_format() == "fp16"


def test_training_gradients_match_reference():
    """Forward + backward MFMA kernels (no update) against the CPU reference
    with the same BF16 rounding points, and the loss they report."""
    model = _plate_model()
    p = ref.init_params(21).astype(np.float32)
    model.surrogate_params(p)
    rng = np.random.default_rng(22)
    x = rng.uniform(-1, 1, size=(512, 7)).astype(np.float32)
    t = rng.normal(scale=0.3, size=(512, 7)).astype(np.float32)
    g, loss = model.surrogate_gradients(x, t)
    lr, gr = ref.loss_and_grad(p.astype(np.float64), x, t, rounded=True)
    assert abs(loss - lr) <= 1e-3 * lr
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in ref.SHAPES])
    for (name, _), a, b in zip(ref.SHAPES, offs[:-1], offs[1:]):
        err = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        assert err < 2e-2, (name, err)
    # and against the exact FP64 gradient, to the BF16 precision
    _, ge = ref.loss_and_grad(p.astype(np.float64), x, t, rounded=False)
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 5e-2


def test_training_set_follows_the_plain_law():
    model = _plate_model()
    model.surrogate_generate(400, 250, seed=3)
    assert model.query("surrogate_samples") == 100000
    x, t = model.surrogate_dataset(0, 100000)
    eps = model.query("surrogate_eps")
    sS, sW = model.query("surrogate_stress_scale"), model.query("surrogate_energy_scale")
    lam, mu = _lame(PLATE["E"], PLATE["nu"])
    E = x[:, :6].astype(np.float64) * eps
    assert np.abs(x[:, :6]).max() <= 1.0 + 1e-6
    assert x[:, 6].min() >= 0 and x[:, 6].max() <= 1
    S, W = ref.neohookean_green(E, lam, mu)
    Sl, Wl = ref.linear_green(E, lam, mu)
    assert np.abs(Sl + t[:, :6] * sS - S).max() <= 1e-5 * np.abs(S).max()
    assert np.abs(Wl + t[:, 6] * sW - W).max() <= 1e-5 * np.abs(W).max()
    # paths reach the envelope and damage grows on some of them
    assert np.abs(x[:, :6]).max() > 0.5
    assert (x[:, 6] > 0).mean() > 0.01


def _stress_errors(model, x, t):
    """Relative error of the surrogate's S_NH and of the linear part alone."""
    eps = model.query("surrogate_eps")
    sS = model.query("surrogate_stress_scale")
    lam, mu = _lame(PLATE["E"], PLATE["nu"])
    E = x[:, :6].astype(np.float64) * eps
    S, _ = ref.neohookean_green(E, lam, mu)
    Sl, _ = ref.linear_green(E, lam, mu)
    y = model.surrogate_infer(x)
    Ss = Sl + y[:, :6].astype(np.float64) * sS
    return _rel(Ss, S), _rel(Sl, S)


def test_training_gains_over_the_linear_law():
    """A short training run: the network's stress is far closer to the
    neo-Hookean stress than what is left without it (its linear part), on
    samples it was not trained on."""
    model = _plate_model()
    model.surrogate_generate(1000, 250, seed=5)
    loss0 = model.surrogate_train(1, batch=4096)
    loss = model.surrogate_train(1500, batch=4096, lr=2e-3)
    assert loss < 0.05 * loss0
    held = _plate_model()
    held.surrogate_generate(40, 250, seed=99)
    x, t = held.surrogate_dataset(0, 10000)
    err, err_lin = _stress_errors(model, x, t)
    print("held-out relative stress error: surrogate %.3e, linear part alone %.3e" % (err, err_lin))
    assert err < 0.1 * err_lin
    assert err < 5e-3


def test_surrogate_step_against_plain_law():
    """The surrogate element kernel inside the explicit step: close to the
    plain mixed step, and with an envelope nothing fits (eps_max tiny), every
    point takes the plain law and the result is the plain step's."""
    def run(enable, eps=0.05, train=True):
        model = _plate_model(eps=eps)
        if train:
            model.surrogate_generate(1000, 250, seed=5)
            model.surrogate_train(1500, batch=4096, lr=2e-3)
        X = model.mesh.nodes
        blocked = np.zeros_like(X, dtype=bool)
        bottom = np.isclose(X[:, 1], 0.0)
        top = np.isclose(X[:, 1], X[:, 1].max())
        blocked[bottom] = True
        model.set("blocked_dofs", blocked)
        dt = 0.8 * model.get_stable_time_step()
        model.set_time_step(dt)
        # 4 m/s: about 3 % mean strain after 300 steps, more at the notch tip
        model.set_dirichlet_increment(np.nonzero(top)[0], 1, 4.0 * dt)
        model.surrogate_enable(enable)
        model.solve_step(300, akx.MIXED)
        return model.get_displacement(), model.get("damage"), model.surrogate_fallbacks()

    u_plain, d_plain, _ = run(False, train=False)
    u_sur, d_sur, fb = run(True)
    print("surrogate vs plain after 300 steps: rel u %.3e, fallbacks %d" % (_rel(u_sur, u_plain), fb))
    # a 1500-iteration network (held-out stress error ~4e-3) moves the
    # displacements by about as much; the full training run does 1e-3
    assert _rel(u_sur, u_plain) < 5e-3
    u_fb, d_fb, fb_all = run(True, eps=1e-9, train=False)
    assert fb_all > 0
    assert _rel(u_fb, u_plain) < 1e-6
    assert np.abs(d_fb - d_plain).max() < 1e-6


def _tension(model, steps, layout, enable):
    X = model.mesh.nodes
    blocked = np.zeros_like(X, dtype=bool)
    blocked[np.isclose(X[:, 1], 0.0)] = True
    top = np.nonzero(np.isclose(X[:, 1], X[:, 1].max()))[0]
    model.set("blocked_dofs", blocked)
    model.set("displacement", np.zeros_like(X))
    model.set("velocity", np.zeros_like(X))
    model.set("acceleration", np.zeros_like(X))
    model.set("damage", np.zeros(model.mesh.nb_elements))
    dt = 0.8 * model.get_stable_time_step()
    model.set_time_step(dt)
    model.set_dirichlet_increment(top, 1, 4.0 * dt)
    model.set_launch(256, 256, 1, 1, layout, 0)
    model.surrogate_enable(enable)
    model.surrogate_fallbacks(reset=True)
    if steps:
        model.solve_step(steps, akx.MIXED)
    return model.get_displacement(), model.get("damage"), model.surrogate_fallbacks()


def test_surrogate_patch_kernel_falls_back_to_the_plain_patch_kernel():
    """Surrogate forces in the LDS patch kernel (slot layouts 2-5): with an
    envelope nothing strained fits (eps_max = 1e-9; unstrained points stay
    inside, where the network's stress, scaled by (lambda + 2 mu) eps_max^2,
    vanishes), every strained point takes the plain FP32 law inside the same
    kernel (the compiler may contract its multiply-adds differently in the
    two instantiations), and the step equals the plain law's patch step
    within 1e-6; the strained points are counted as fallbacks."""
    model = _plate_model(eps=1e-9)
    if model.query("patch_available") != 1:
        pytest.skip("no LDS patches on this mesh")
    E = model.mesh.nb_elements
    for layout in (2, 4, 5):
        u_plain, d_plain, _ = _tension(model, 150, layout, False)
        u_sur, d_sur, fb = _tension(model, 150, layout, True)
        assert _rel(u_sur, u_plain) < 1e-6, layout
        assert np.abs(d_sur - d_plain).max() < 1e-6, layout
        assert 0.5 * 150 * E < fb <= 150 * E


def test_surrogate_patch_kernel_matches_element_kernel():
    """With a trained network, the surrogate step through the patch kernel
    (per-patch FP64 partial sums) and through k_sur_forces (FP32 element
    slots, layout 0) apply the same network: displacements within 1e-5 of
    each other after 300 steps, damage within 5e-4 (the network's FP16
    roundings make its output jump by ~1e-3 of the remainder where a
    rounding flips, and the damage update amplifies the energy's change
    near a threshold)."""
    model = _plate_model()
    model.surrogate_generate(400, 250, seed=5)
    model.surrogate_train(500, batch=4096, lr=2e-3)
    u0, d0, fb0 = _tension(model, 300, 0, True)
    u5, d5, fb5 = _tension(model, 300, 5, True)
    assert _rel(u5, u0) < 1e-5
    assert np.abs(d5 - d0).max() < 5e-4
    assert abs(fb0 - fb5) <= 1e-3 * fb0


@pytest.mark.parametrize("layout", [0, 5])
def test_surrogate_step_in_a_graph(layout):
    """The surrogate step (strain pass, network, forces, node update)
    captured in a HIP graph and replayed four times, with eager reads of the
    state between the replays, equals the same four runs launched eagerly,
    bit for bit (displacements, damage, fallback count)."""
    model = _plate_model()
    model.surrogate_generate(200, 250, seed=5)
    model.surrogate_train(200, batch=4096, lr=2e-3)
    out = {}
    for name in ("eager", "graph"):
        _tension(model, 0, layout, True)
        us = []
        for _ in range(4):
            if name == "eager":
                model.solve_step(10, akx.MIXED)
            else:
                model.solve_step_graph(10, akx.MIXED)
            us.append(model.get_displacement())
        out[name] = (us, model.get("damage"), model.surrogate_fallbacks())
    for ue, ug in zip(out["eager"][0], out["graph"][0]):
        assert np.array_equal(ue, ug)
    assert np.array_equal(out["eager"][1], out["graph"][1])
    assert out["eager"][2] == out["graph"][2]
    assert np.abs(out["eager"][0][-1]).max() > 0


def test_surrogate_step_bf16_format():
    """The surrogate step with the network in BF16: a graph captured in FP16
    is not replayed after the switch (the graph step equals the eager BF16
    step bit for bit), and the two formats of one trained network give
    displacements within 2e-3 of each other, not equal."""
    model = _plate_model()
    model.surrogate_generate(200, 250, seed=5)
    model.surrogate_train(200, batch=4096, lr=2e-3)
    u16 = _tension(model, 100, 5, True)[0]
    _tension(model, 0, 5, True)
    model.solve_step_graph(100, akx.MIXED)
    assert np.array_equal(model.get_displacement(), u16)
    model.surrogate_format("bf16")
    ub = _tension(model, 100, 5, True)[0]
    _tension(model, 0, 5, True)
    model.solve_step_graph(100, akx.MIXED)
    assert np.array_equal(model.get_displacement(), ub)
    print("BF16 vs FP16 network after 100 steps: rel u %.3e" % _rel(ub, u16))
    assert 0 < _rel(ub, u16) < 2e-3


def test_surrogate_on_one_of_two_materials():
    """The surrogate bound to the neohookean_damage material of a two-material
    plate (the other elastic): its kernels then run on that material's
    element list (a filtered launch group). With an envelope no strained
    point fits, the step equals the plain law's within 1e-6; with a trained
    network it stays within 5e-3 of it, and the elastic part is untouched by
    the switch (same law on both runs)."""
    m = akm.notched_plate_tet4(12, 12, 3)
    E = m.nb_elements
    X = m.nodes
    cent = X[m.conn].mean(axis=1)
    em = (cent[:, 0] > 0.5 * X[:, 0].max()).astype(np.int32)  # right half: damage law
    yd = akrng.weibull_thresholds(E, 1e6, 5.0, seed=4)

    def model(eps):
        g = akx.SolidMechanicsModel(m)
        g.add_material(E=PLATE["E"], nu=PLATE["nu"], rho=PLATE["rho"])
        g.add_material(kind=akx.NEOHOOKEAN_DAMAGE, Sd=SD, Yd=yd, **PLATE)
        g.assign_materials(em)
        g.init_full()
        g.surrogate_setup(1, eps)
        return g

    def run(g, enable):
        blocked = np.zeros_like(X, dtype=bool)
        blocked[np.isclose(X[:, 1], 0.0)] = True
        top = np.nonzero(np.isclose(X[:, 1], X[:, 1].max()))[0]
        g.set("blocked_dofs", blocked)
        dt = 0.8 * g.get_stable_time_step()
        g.set_time_step(dt)
        g.set_dirichlet_increment(top, 1, 4.0 * dt)
        g.surrogate_enable(enable)
        g.solve_step(200, akx.MIXED)
        return g.get_displacement(), g.get("damage")

    u_plain, d_plain = run(model(0.05), False)
    g = model(1e-9)
    u_fb, d_fb = run(g, True)
    assert g.surrogate_fallbacks() > 0
    assert _rel(u_fb, u_plain) < 1e-6
    assert np.abs(d_fb - d_plain).max() < 1e-6
    g = model(0.05)
    g.surrogate_generate(1000, 250, seed=5)
    g.surrogate_train(1500, batch=4096, lr=2e-3)
    u_sur, d_sur = run(g, True)
    assert _rel(u_sur, u_plain) < 5e-3
    assert (d_sur[em == 0] == 0).all() and (d_plain[em == 0] == 0).all()
