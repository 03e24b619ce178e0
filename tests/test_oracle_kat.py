"""Pins the CPU FP64 oracle against the reference's own known-answer tests.

Each test restates one reference test (cited) on the reference's own fixture
files (copied under tests/golden/), or checks a property the mathematics fixes
exactly.  No GPU is used.
"""
import os

import numpy as np
import pytest

from akantu_amd import mesh as akm
from oracle import oracle as orc
from tests import patch_helpers as ph

GOLDEN = ph.GOLDEN


def test_truss_dynamic_explicit_verified():
    """test/test_model/test_common/test_model_solver/test_model_solver_dynamic.cc
    (EXPLICIT=true) must print test_model_solver_dynamic_explicit.verified.
    Pins predictor / residual / lumped solve / corrector and their order."""
    ref = open(os.path.join(GOLDEN, "test_model_solver_dynamic_explicit.verified")).read()
    out = orc.truss_dynamic_explicit()
    assert out.splitlines() == ref.splitlines()


def _elastic_model(m, **kw):
    mat = dict(ph.STEEL)
    mat.update(kw)
    return orc.OracleModel(m, [mat])


@pytest.mark.parametrize("name", ph.PATCH_TYPES)
def test_gradient_and_integrate_kat(name):
    """test/test_fe_engine/test_gradient.cc:46-102 and test_integrate.cc:41-72 on the
    reference's fe-engine fixture meshes: gradient of a linear field is exact
    (5e-13), gradient of the coordinates is the identity (2e-14), the
    integral of a constant equals the (unit) mesh volume (1e-14)."""
    m = akm.read_msh(os.path.join(GOLDEN, "fe_%s.msh" % name))
    o = _elastic_model(m)
    d = m.dim
    # integrate 1 -> volume 1
    assert abs(o.field("JxW").sum() - 1.0) < 1e-14
    # gradient of alpha . x via the displacement slot
    alpha = np.array([[13, 23, 31], [11, 7, 5], [2, 3, 4]], dtype=float)[:d, :d]
    o.nodal("u")[:] = m.nodes @ alpha.T
    o.assemble_internal_forces()
    G = o.field("gradu").reshape(-1, d, d).transpose(0, 2, 1)  # col-major storage
    assert np.abs(G - alpha[None]).max() < 5e-13
    o.nodal("u")[:] = m.nodes
    o.assemble_internal_forces()
    G = o.field("gradu").reshape(-1, d, d).transpose(0, 2, 1)
    assert np.abs(G - np.eye(d)[None]).max() < 2e-14


@pytest.mark.parametrize("name", ph.PATCH_TYPES)
def test_lumped_mass_sums_to_mass(name):
    """test/test_model/test_solid_mechanics_model/test_solid_mechanics_model_mass/
    test_lumped_mass.cc:81-101: the lumped mass of every direction sums to
    rho * volume (2e-15 relative)."""
    m = akm.read_msh(os.path.join(GOLDEN, "fe_%s.msh" % name))
    o = _elastic_model(m, rho=1.0)
    M = o.nodal("mass")
    vol = o.field("JxW").sum()
    for k in range(m.dim):
        assert abs(M[:, k].sum() - vol) / vol < 2e-15
    assert np.all(M[:, 0] > 0)


@pytest.mark.parametrize("name", ph.PATCH_TYPES)
def test_patch_linear_elastic_explicit(name):
    """patch_test_linear_elastic_explicit.cc:33-68 with the fixture of
    patch_test_linear_fixture.hh: boundary nodes blocked, all nodes on the
    static linear solution, 100 explicit steps at 0.8 x stable dt; then
    E_kin = 0 (1e-16), DOFs exact (1e-15), grad u and stress exact (1e-13
    relative), sum f_int = 0 and f_int + f_ext = 0 (1e-9 of max |f_int|)."""
    m = ph.patch_mesh(name)
    d = m.dim
    o = _elastic_model(m)
    bnodes = akm.boundary_nodes(m)
    o.blocked[bnodes, :] = 1
    o.nodal("u")[:] = ph.linear_field(m.nodes)
    o.set_time_step(0.8 * o.stable_time_step())
    o.solve_step(100)
    assert abs(o.energy("kinetic")) < 1e-16
    assert np.abs(o.nodal("u") - ph.linear_field(m.nodes)).max() < 1e-15
    pg = ph.prescribed_gradient(d)
    G = o.field("gradu").reshape(-1, d, d).transpose(0, 2, 1)
    gerr = np.abs(G - pg[None]).max(axis=(1, 2)) / np.abs(G).max(axis=(1, 2))
    assert gerr.max() < 1e-13
    ps = ph.plane_strain_stress(pg)
    S = o.field("stress").reshape(-1, d, d).transpose(0, 2, 1)
    serr = np.abs(S - ps[None]).max(axis=(1, 2)) / np.abs(ps).max()
    assert serr.max() < 1e-13
    fint = o.nodal("f_int")
    fmax = np.abs(fint).max()
    assert np.abs(fint.sum(axis=0)).max() / fmax < 1e-9
    fext = ph.boundary_traction_forces(m, S[0])
    assert np.abs(fint + fext).max() / fmax < 1e-9


def _random_rotation(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@pytest.mark.parametrize("d", [2, 3])
def test_elastic_stress_rotated_composed_strain(d):
    """test_elastic_materials.cc:109-136 (2D, 1e-13) and :213-240 (3D, 1e-14):
    stress of a rotated composed strain, rotated back, equals the K/mu closed
    form (E = 1, nu = .3)."""
    E, nu = 1.0, 0.3
    rng = np.random.default_rng(0)
    dev = np.array([[0., 1., 2.], [1., 0., 3.], [2., 3., 0.]])
    grad = (np.eye(3) + dev)[:d, :d]
    K = E / (3 * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    for _ in range(5):
        R = _random_rotation(d, rng)
        grad_rot = R @ (grad @ R.T)
        S_rot, _ = orc.point_stress(orc.ELASTIC, d, E, nu, grad_rot)
        S = R.T @ (S_rot @ R)
        strain = 0.5 * (grad + grad.T)
        dev_strain = strain - np.trace(strain) / 3.0 * np.eye(d)
        expected = 2 * mu * dev_strain + (d * K) * np.eye(d)
        err = np.abs(S - expected).max()
        if d == 2:
            err /= np.abs(expected).max()
        assert err < (1e-13 if d == 2 else 1e-14)


def test_bar_energy_conservation_and_order():
    """Central difference conserves E_kin + E_pot to O(dt^2) for a linear
    elastic bar with a free end; a factor error in the kinetic or potential
    energy, in the mass or in the force sign would break it."""
    m = akm.bar_tri3(40, 8)
    o = _elastic_model(m, E=210e9, nu=0.3, rho=7800.0)
    X = m.nodes
    left = np.isclose(X[:, 0], 0.0)
    o.blocked[left, :] = 1
    o.nodal("v")[:, 0] = 0.01 * np.sin(np.pi * X[:, 0] / 2)
    o.set_time_step(0.8 * o.stable_time_step())
    e0 = o.energy("kinetic") + o.energy("potential")
    worst = 0.0
    for _ in range(50):
        o.solve_step(10)
        e = o.energy("kinetic") + o.energy("potential")
        worst = max(worst, abs(e - e0) / e0)
    assert worst < 5e-3
    assert o.energy("potential") > 0


def test_stable_time_step_closed_form():
    """c0 bar: h = 2 r_in of the right triangle with legs 0.01 (inradius
    (a + b - c)/2), c_p = sqrt((lambda + 2 mu)/rho) in plane strain."""
    m = akm.bar_tri3()
    o = _elastic_model(m, E=210e9, nu=0.3, rho=7800.0)
    a = 0.01
    r = (a + a - np.sqrt(2) * a) / 2
    E, nu, rho = 210e9, 0.3, 7800.0
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    cp = np.sqrt((lam + 2 * mu) / rho)
    assert abs(o.stable_time_step() - 2 * r / cp) / (2 * r / cp) < 1e-12
