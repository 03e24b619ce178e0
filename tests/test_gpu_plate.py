"""The notched plate under tension with the neo-Hookean + Marigo damage law on
tetrahedron_4 (the learned-surrogate workload, reduced), against the oracle.

Law: S = (1 - d) S_NH(F) with the Marigo update on the driving energy
Y = W_NH (material_neohookean_inline_impl.hh:46-65,145-158,
material_marigo_inline_impl.hh:67-76); per-element thresholds Yd from a
Weibull(m=5, scale 0.2 MPa) draw of the reference's engine (seed 4), Sd = 20 MPa.  The top face is pulled at a
prescribed velocity (applyBC(IncrementValue) before every solveStep), the
bottom face is held: after 300 steps about 3/4 of the points are damaged.

* FP64: displacements and damage within 1e-12 of the oracle, the damage
  branch (d > 0) equal on every point.
* Mixed (FP32 element law), every slot layout and table mode the autotuner
  can pick: displacements within 1e-5 after 300 steps; the share of points
  whose damaged / undamaged state differs from the oracle and the largest
  damage difference, each under a stated bound.
* The surrogate with an input envelope no point fits: every point takes the
  plain law, and the step is within 1e-5 of the oracle.
"""
import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from akantu_amd import rng as akrng
from tests import gpu_helpers as gh

pytestmark = pytest.mark.gpu

PLATE = dict(kind=akx.NEOHOOKEAN_DAMAGE, E=1e9, nu=0.3, rho=1200.0)
SD = 2e7
SPEED = 4.0        # m/s on the top face: ~3 % mean strain after 300 steps
STEPS = 300


def plate_pair(nx=12, ny=12, nz=3, seed=4, yd_scale=2e5):
    m = akm.notched_plate_tet4(nx, ny, nz)
    yd = akrng.weibull_thresholds(m.nb_elements, yd_scale, 5.0, seed=seed)
    p = gh.Pair(m, [dict(PLATE, Sd=SD, Yd=yd)])
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 1], 0.0)] = True
    top = np.where(np.isclose(X[:, 1], X[:, 1].max()))[0]
    mask[top, 1] = True
    p.block(mask)
    dt = 0.8 * p.cpu.stable_time_step()
    p.set_time_step(dt)
    p.top, p.inc = top, SPEED * dt
    p.gpu.set_dirichlet_increment(top, 1, p.inc)
    return p


def oracle_steps(p, nsteps):
    u = p.cpu.nodal("u")
    for _ in range(nsteps):
        u[p.top, 1] += p.inc
        p.cpu.solve_step(1)


@pytest.fixture(scope="module")
def oracle_run():
    p = plate_pair()
    oracle_steps(p, STEPS)
    u = p.cpu.nodal("u").copy()
    d = p.cpu.field("damage").copy()
    # about 3/4 of the points damaged, none broken (d ~ 0.12 at most)
    assert (d > 0).mean() > 0.5, "the load must damage a fair share of the points"
    assert (d < 0.5).all()
    return u, d, p.cpu.energy("potential")


def test_plate_damage_fp64_matches_oracle(oracle_run):
    u_ref, d_ref, ep_ref = oracle_run
    p = plate_pair()
    p.gpu.solve_step(STEPS, akx.FP64)
    u, d = p.gpu.get_displacement(), p.gpu.get("damage")
    assert gh.rel(u, u_ref) <= 1e-12
    assert np.array_equal(d > 0, d_ref > 0)
    assert np.abs(d - d_ref).max() <= 1e-12
    ep = p.gpu.get_energy("potential")
    assert abs(ep - ep_ref) <= 1e-12 * abs(ep_ref)


@pytest.mark.parametrize("recompute,layout", [(0, 0), (1, 0), (1, 1), (1, 2), (1, 3), (1, 4),
                                              (1, 5), (1, 6)])
def test_plate_damage_mixed_matches_oracle(oracle_run, recompute, layout):
    """Layout 6 has no fused form for a law with history and runs as 5."""
    u_ref, d_ref, _ = oracle_run
    p = plate_pair()
    p.gpu.set_launch(mixed_recompute=recompute, slot_layout=layout)
    p.gpu.solve_step(STEPS, akx.MIXED)
    u, d = p.gpu.get_displacement(), p.gpu.get("damage")
    err = gh.rel(u, u_ref)
    flipped = float(np.mean((d > 0) != (d_ref > 0)))
    dmax = float(np.abs(d - d_ref).max())
    print("layout %d recompute %d: rel u %.3e, branch disagreement %.4f %%, max |dd| %.3e"
          % (layout, recompute, err, 100 * flipped, dmax))
    assert err <= 1e-5
    # d = (Y - Yd)/Sd is continuous at the threshold, so a point whose
    # branch differs sits at Y ~ Yd and carries a damage near zero
    assert flipped <= 0.005
    assert dmax <= 1e-5


def test_plate_surrogate_fallback_matches_oracle(oracle_run):
    """An envelope of |E_k| <= 1e-9: the surrogate kernels send every point
    to the plain law (FP32), which then tracks the oracle as the plain mixed
    step does."""
    u_ref, d_ref, _ = oracle_run
    p = plate_pair()
    p.gpu.surrogate_setup(0, 1e-9)
    p.gpu.surrogate_enable(True)
    p.gpu.solve_step(STEPS, akx.MIXED)
    u, d = p.gpu.get_displacement(), p.gpu.get("damage")
    fb = p.gpu.surrogate_fallbacks()
    # points whose strain is exactly zero (ahead of the wave) or below 1e-9
    # stay inside the envelope; the others are counted once per step
    assert fb >= 0.5 * p.mesh.nb_elements * STEPS
    assert gh.rel(u, u_ref) <= 1e-5
    assert float(np.mean((d > 0) != (d_ref > 0))) <= 0.005
    assert np.abs(d - d_ref).max() <= 1e-5
