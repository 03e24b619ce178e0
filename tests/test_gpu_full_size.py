"""The two 3D workloads at their full sizes, checked through properties that do
not depend on the mesh size:

* the 6,000,000-element tetrahedron_4 cube (100^3 hexes, 6 tets each,
  interior jitter, bottom fixed, top face at -1 m/s, dt = 0.8 x stable):
  FP64 == the oracle bit for bit over the first steps; mixed within 1e-5 of
  FP64 after 200 more steps from the same state; the internal force of the
  elastic law is linear, and scaling u by 2 scales every rounded operation
  exactly, so f(2u) == 2 f(u) bit for bit on both paths; a rigid translation
  makes every mixed-path difference u_a - u_ref exactly zero, hence zero
  forces;
* the 128^3 hexahedron_8 neo-Hookean cube (bottom fixed, top face pushed at
  0.5 m/s): mixed within 1e-5 of FP64 after 100 steps, and a rigid
  translation gives zero mixed forces;
* the 4.03 M-element notched plate with the neo-Hookean + damage law: zero
  forces under a rigid translation, with and without the surrogate kernels;
  mixed and surrogate-fallback forces within 1e-5 of FP64; mixed tension
  run within 1e-5 of FP64 with the same damaged points.

The oracle runs on the host for the bitwise check only (about 0.3 s per step
at this size)."""
import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from akantu_amd import rng as akrng
from tests import gpu_helpers as gh

pytestmark = pytest.mark.gpu


def _cube_gpu(mesh):
    """The cube set-up of gpu_helpers.cube_pair, GPU model only."""
    g = akx.SolidMechanicsModel(mesh)
    g.add_material(E=210e9, nu=0.3, rho=7800.0)
    g.init_full()
    X = mesh.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 2], 0.0)] = True
    g.set("blocked_dofs", mask)
    g.set_time_step(0.8 * g.get_stable_time_step())
    return g


def _smooth_field(X, scale):
    return scale * np.stack([np.sin(2.0 * X[:, 1] + X[:, 2]),
                             np.cos(1.5 * X[:, 0]) * X[:, 2],
                             X[:, 0] * X[:, 1]], axis=1)


def test_cube_tet4_full_size():
    p = gh.cube_pair(100)
    X = p.mesh.nodes
    assert p.mesh.conn.shape[0] == 6_000_000
    # FP64 == oracle over the first steps
    p.step(4, akx.FP64)
    assert np.array_equal(p.gpu.get_displacement(), p.cpu.nodal("u"))
    assert np.array_equal(p.gpu.get_velocity(), p.cpu.nodal("v"))
    del p.cpu
    # mixed vs FP64 from the same state, in a second GPU model
    mixed = _cube_gpu(p.mesh)
    for name in ("displacement", "velocity", "acceleration"):
        mixed.set(name, p.gpu.get(name))
    p.gpu.solve_step(200, akx.FP64)
    mixed.solve_step(200, akx.MIXED)
    u64 = p.gpu.get_displacement()
    assert np.abs(u64).max() > 0
    assert gh.rel(mixed.get_displacement(), u64) <= 1e-5
    del mixed
    # linearity: f(2u) == 2 f(u) bit for bit, FP64 and mixed
    u = _smooth_field(X, 1e-4)
    for precision in (akx.FP64, akx.MIXED):
        p.gpu.set("displacement", u)
        p.gpu.assemble_internal_forces(precision)
        f1 = p.gpu.get_internal_force()
        p.gpu.set("displacement", 2.0 * u)
        p.gpu.assemble_internal_forces(precision)
        f2 = p.gpu.get_internal_force()
        assert np.abs(f1).max() > 0
        assert np.array_equal(f2, 2.0 * f1), precision
    # rigid translation: exactly zero forces on the mixed path
    p.gpu.set("displacement", np.tile([3e-3, -2e-3, 1e-3], (X.shape[0], 1)))
    p.gpu.assemble_internal_forces(akx.MIXED)
    assert np.all(p.gpu.get_internal_force() == 0.0)


def _hex_gpu(mesh):
    """The set-up of gpu_helpers.hex_pair, GPU model only."""
    g = akx.SolidMechanicsModel(mesh)
    g.add_material(kind=akx.NEOHOOKEAN, E=1e7, nu=0.45, rho=1000.0)
    g.init_full()
    X = mesh.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 2], 0.0)] = True
    g.set("blocked_dofs", mask)
    dt = 0.8 * g.get_stable_time_step()
    g.set_time_step(dt)
    g.set_dirichlet_increment(np.where(np.isclose(X[:, 2], 1.0))[0], 2, -0.5 * dt)
    return g


def test_hex8_neohookean_full_size():
    mesh = akm.cube_hex8(128)
    runs = {}
    for precision in (akx.FP64, akx.MIXED):
        g = _hex_gpu(mesh)
        g.solve_step(100, precision)
        runs[precision] = g.get_displacement()
    assert np.abs(runs[akx.FP64]).max() > 0
    assert gh.rel(runs[akx.MIXED], runs[akx.FP64]) <= 1e-5
    g.set("displacement", np.tile([1e-3, 2e-3, -1e-3], (mesh.nodes.shape[0], 1)))
    g.assemble_internal_forces(akx.MIXED)
    assert np.all(g.get_internal_force() == 0.0)


def _plate_gpu(mesh, yd, surrogate_eps=None):
    """The notched plate at its full size (150 x 150 x 30 cells, 4.03 M
    tetrahedron_4) with the neo-Hookean + damage law, GPU model only."""
    g = akx.SolidMechanicsModel(mesh)
    g.add_material(kind=akx.NEOHOOKEAN_DAMAGE, E=1e9, nu=0.3, rho=1200.0, Sd=2e7, Yd=yd)
    g.init_full()
    X = mesh.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 1], 0.0)] = True
    top = np.where(np.isclose(X[:, 1], X[:, 1].max()))[0]
    mask[top, 1] = True
    g.set("blocked_dofs", mask)
    dt = 0.8 * g.get_stable_time_step()
    g.set_time_step(dt)
    g.set_dirichlet_increment(top, 1, 4.0 * dt)
    if surrogate_eps is not None:
        g.surrogate_setup(0, surrogate_eps)
        g.surrogate_enable(True)
    return g


def test_plate_damage_full_size():
    """4.03 M tetrahedron_4 notched plate, neo-Hookean + damage:
    * a rigid translation gives exactly zero mixed-path forces and no damage,
      with the plain law and with the surrogate's kernels: a zero strain is
      inside any envelope, and the network as initialised (zero biases)
      maps a zero input to a zero output;
    * under a smooth displacement field the mixed forces, and those of the
      surrogate kernels with an envelope no strained point fits (all fall
      back to the plain law), are within 1e-5 of the FP64 forces;
    * 150 steps of tension: mixed within 1e-5 of FP64, the damaged /
      undamaged state equal on at least 99.5 % of the points."""
    mesh = akm.notched_plate_tet4()
    E = mesh.nb_elements
    assert E > 4_000_000
    X = mesh.nodes
    yd = akrng.weibull_thresholds(E, 2e5, 5.0, seed=4)
    shift = np.tile([3e-3, -2e-3, 1e-3], (X.shape[0], 1))
    u = 2e-3 * np.stack([np.sin(20.0 * X[:, 1] + 30.0 * X[:, 2]),
                         np.cos(15.0 * X[:, 0]) * X[:, 1] * 10.0,
                         X[:, 0] * X[:, 2] * 50.0], axis=1)
    forces = {}
    for name, precision, eps in (("fp64", akx.FP64, None), ("mixed", akx.MIXED, None),
                                 ("surrogate", akx.MIXED, 1e-9)):
        g = _plate_gpu(mesh, yd, eps)
        if precision == akx.MIXED:
            g.set("displacement", shift)
            g.assemble_internal_forces(precision)
            assert np.all(g.get_internal_force() == 0.0), name
            assert np.all(g.get("damage") == 0.0), name
            if eps is not None:
                assert g.surrogate_fallbacks(reset=True) == 0
        g.set("displacement", u)
        g.assemble_internal_forces(precision)
        forces[name] = g.get_internal_force()
        if eps is not None:
            assert g.surrogate_fallbacks() >= 0.99 * E
        del g
    assert np.abs(forces["fp64"]).max() > 0
    for name in ("mixed", "surrogate"):
        assert gh.rel(forces[name], forces["fp64"]) <= 1e-5, name
    runs = {}
    for precision in (akx.FP64, akx.MIXED):
        g = _plate_gpu(mesh, yd)
        g.solve_step(150, precision)
        runs[precision] = (g.get_displacement(), g.get("damage"))
        del g
    (u64, d64), (u32, d32) = runs[akx.FP64], runs[akx.MIXED]
    assert (d64 > 0).any()
    assert gh.rel(u32, u64) <= 1e-5
    assert float(np.mean((d32 > 0) != (d64 > 0))) <= 0.005
    assert np.abs(d32 - d64).max() <= 1e-5
