"""Pins of the oracle's FEEngine and Material operators on caller fields
(fe_engine.hh:95-244, material.hh:110-195), on the reference's own test
meshes (test/test_fe_engine/_*.msh):

* gradientOnIntegrationPoints of a linear field is its constant gradient
  (test_gradient.cc:46-102), for 1 to 3 components, with and without an
  element filter;
* interpolateOnIntegrationPoints reproduces a linear field at the points
  (partition of unity and linear completeness of N);
* integrate of 1 is the mesh volume, of x the first moment
  (test_integrate.cc:41-72); the scalar form is Math::reduce of the element
  integrals; integrateOnIntegrationPoints summed over the points is
  integrate;
* computeBtD is the product D·B of each point (checked with numpy), computeNtb
  the outer product b N_a;
* assembleFieldLumped of rho is the model's lumped mass, bit for bit;
* Material::computeStress equals the pinned point law; computeAllStresses +
  assembleInternalForces, material by material, equal the whole-model
  assembly bit for bit; the materials' potential energies add up to the
  model's.
"""
import os

import numpy as np
import pytest

from akantu_amd import mesh as akm
from oracle import oracle as orc

TYPES = ["triangle_3", "quadrangle_4", "tetrahedron_4", "hexahedron_8"]


def _mesh(golden_dir, name):
    return akm.read_msh(os.path.join(golden_dir, "fe_%s.msh" % name))


def _model(mesh, mats=None, elem_mat=None):
    mats = mats or [dict(E=1.0, nu=0.3, rho=2.5)]
    return orc.OracleModel(mesh, mats, elem_mat)


def _volume(mesh):
    return 1.0 if mesh.dim == 3 else 1.0


@pytest.mark.parametrize("name", TYPES)
def test_gradient_of_linear_field(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    M = _model(mesh)
    d, Q = mesh.dim, mesh.nb_qp
    rng = np.random.default_rng(3)
    for nb in (1, 2, 3):
        A = rng.uniform(-1, 1, size=(nb, d))
        c = rng.uniform(-1, 1, size=nb)
        u = mesh.nodes @ A.T + c
        g = M.fe_gradient(u, nb)
        assert g.shape == (mesh.nb_elements * Q, nb * d)
        expect = A.T.reshape(-1)      # column-major nb x d: [j*nb + i] = A[i, j]
        assert np.abs(g - expect).max() <= 5e-13
        ids = rng.choice(mesh.nb_elements, size=min(7, mesh.nb_elements),
                         replace=False).astype(np.uint32)
        gf = M.fe_gradient(u, nb, ids)
        rows = (ids[:, None] * Q + np.arange(Q)).reshape(-1)
        assert np.array_equal(gf, g[rows])


@pytest.mark.parametrize("name", TYPES)
def test_interpolation_is_linear_complete(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    M = _model(mesh)
    ones = M.fe_interpolate(np.ones(mesh.nb_nodes), 1)
    assert np.abs(ones - 1.0).max() <= 1e-15
    xq = M.fe_interpolate(mesh.nodes, mesh.dim)
    A = np.array([[0.3, -1.2, 0.7][:mesh.dim], [2.0, 0.5, -0.4][:mesh.dim]])
    uq = M.fe_interpolate(mesh.nodes @ A.T + 1.5, 2)
    assert np.abs(uq - (xq @ A.T + 1.5)).max() <= 1e-14
    # the points lie inside their elements: on each element's bounding box
    conn = mesh.conn.astype(np.int64)
    lo = mesh.nodes[conn].min(axis=1).repeat(mesh.nb_qp, axis=0)
    hi = mesh.nodes[conn].max(axis=1).repeat(mesh.nb_qp, axis=0)
    assert np.all(xq >= lo - 1e-14) and np.all(xq <= hi + 1e-14)


@pytest.mark.parametrize("name", TYPES)
def test_integrals(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    M = _model(mesh)
    Q, E = mesh.nb_qp, mesh.nb_elements
    one = np.ones((E * Q, 1))
    per = M.fe_integrate(one, 1)
    assert per.shape == (E, 1)
    vol = per.sum()
    assert vol == pytest.approx(_volume(mesh), abs=1e-14)
    xq = M.fe_interpolate(mesh.nodes, mesh.dim)
    first = M.fe_integrate(xq, mesh.dim).sum(axis=0)
    # the test meshes are unit squares / cubes centred on 0.5 or 0
    centre = first / vol
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    assert np.abs(centre - 0.5 * (lo + hi)).max() <= 1e-13
    # the scalar form: Math::reduce (sorted, pairwise) of the element integrals
    f = np.sin(3.0 * xq[:, 0]) + xq[:, 1] ** 2
    vals = M.fe_integrate(f, 1)[:, 0]
    assert M.fe_integrate_scalar(f) == pytest.approx(vals.sum(), rel=1e-14)
    onp = M.fe_integrate_on_points(np.stack([f, 2 * f], axis=1), 2)
    summed = onp.reshape(E, Q, 2).sum(axis=1)
    assert np.abs(summed[:, 0] - vals).max() <= 1e-15 * max(1.0, np.abs(vals).max()) * Q
    assert np.array_equal(summed[:, 1], 2 * summed[:, 0])


def test_math_reduce_order():
    """The scalar integral is the reference's Math::reduce: sorted ascending,
    then summed pairwise (an odd tail folded into its neighbour first) -- a
    result different in the last bits from a plain left-to-right sum."""
    m = akm.rectangle_tri3(5, 1, 1.0, 1.0)   # 10 elements, all of area 0.1
    M = _model(m)
    rng = np.random.default_rng(1)
    f = rng.standard_normal(10) * 10.0 ** rng.integers(-8, 8, size=10)
    per = M.fe_integrate(f, 1)[:, 0]
    a = list(np.sort(per))
    n = len(a)
    half = n >> 1
    while half:
        if n - 2 * half:
            a[n - 2] += a[n - 1]
        for i in range(half):
            a[i] = a[2 * i] + a[2 * i + 1]
        n, half = half, half >> 1
    assert M.fe_integrate_scalar(f) == a[0]


@pytest.mark.parametrize("name", TYPES)
def test_btd_and_ntb_products(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    M = _model(mesh)
    d, n, Q, E = mesh.dim, mesh.conn.shape[1], mesh.nb_qp, mesh.nb_elements
    rng = np.random.default_rng(8)
    B = M.field("B").reshape(E * Q, n, d).transpose(0, 2, 1)      # (points, d, n)
    for rows in (1, d):
        D = rng.standard_normal((E * Q, rows * d))
        out = M.fe_btd(D, rows)
        Dm = D.reshape(E * Q, d, rows).transpose(0, 2, 1)          # column-major rows x d
        expect = np.einsum("pij,pjk->pik", Dm, B)                  # (points, rows, n)
        got = out.reshape(E * Q, n, rows).transpose(0, 2, 1)
        assert np.abs(got - expect).max() <= 1e-13 * np.abs(expect).max()
    b = rng.standard_normal((E * Q, 2))
    ntb = M.fe_ntb(b, 2).reshape(E * Q, n, 2)
    # N_a at the points: interpolation of the indicator of element 0's node a
    Ntab = np.zeros((Q, n))
    for a in range(n):
        ind = np.zeros(mesh.nb_nodes)
        ind[mesh.conn[0, a]] = 1.0
        Ntab[:, a] = M.fe_interpolate(ind, 1, np.array([0], dtype=np.uint32))[:, 0]
    assert np.abs(Ntab.sum(axis=1) - 1.0).max() <= 1e-15
    Na = np.tile(Ntab, (E, 1))
    assert np.abs(ntb - Na[:, :, None] * b[:, None, :]).max() <= 1e-15 * np.abs(b).max()
    assert np.abs(ntb.sum(axis=1) - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("name", TYPES)
def test_lumped_assembly_is_the_mass(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    rho = 2.5
    M = _model(mesh, [dict(E=1.0, nu=0.3, rho=rho)])
    field = np.full((mesh.nb_elements * mesh.nb_qp, 1), rho)
    lumped = M.fe_assemble_field_lumped(field, 1, np.zeros((mesh.nb_nodes, 1)))
    assert np.array_equal(lumped[:, 0], M.nodal("mass")[:, 0])
    # added to what is there, one addition per node
    twice = M.fe_assemble_field_lumped(field, 1, lumped)
    assert np.array_equal(twice, lumped + lumped)


@pytest.mark.parametrize("name", TYPES)
def test_material_operators(golden_dir, name):
    mesh = _mesh(golden_dir, name)
    d, Q, E = mesh.dim, mesh.nb_qp, mesh.nb_elements
    em = (np.arange(E) % 3 == 0).astype(np.int32)
    mats = [dict(kind=orc.ELASTIC, E=210e9, nu=0.3, rho=7800.0),
            dict(kind=orc.NEOHOOKEAN, E=10e6, nu=0.45, rho=1000.0)]
    M = _model(mesh, mats, em)
    rng = np.random.default_rng(4)
    u = 1e-3 * rng.standard_normal(mesh.nodes.shape)
    M.nodal("u")[:] = u
    # whole-model assembly
    M.assemble_internal_forces()
    f_all = M.nodal("f_int").copy()
    # material by material on a cleared internal force
    M.nodal("f_int")[:] = 0.0
    for k in range(2):
        M.material_compute_all_stresses(k)
    for k in range(2):
        M.material_assemble_internal_forces(k)
    assert np.array_equal(M.nodal("f_int"), f_all)
    # computeStress from given gradients equals the point law
    for k, mat in enumerate(mats):
        ids = M.material_elements(k)
        assert np.array_equal(ids, np.nonzero(em == k)[0])
        G = 1e-2 * rng.standard_normal((len(ids) * Q, d * d))
        S = M.material_compute_stress(k, G)
        for p in range(0, len(G), max(1, len(G) // 5)):
            Gm = G[p].reshape(d, d).T
            Sp, _ = orc.point_stress(mat["kind"], d, mat["E"], mat["nu"], Gm)
            assert np.array_equal(S[p].reshape(d, d).T, Sp)
    # potential energies of the materials add up to the model's
    ep = [M.material_potential_energy(k) for k in range(2)]
    assert ep[0] + ep[1] == M.energy("potential")


def test_damage_material_compute_stress_updates_history():
    m = akm.box_tet4(2, 2, 2)
    Q = m.nb_qp
    yd = np.full(m.nb_elements * Q, 10.0)
    mat = dict(kind=orc.NEOHOOKEAN_DAMAGE, E=1e4, nu=0.3, rho=1.0, Sd=100.0, Yd=yd)
    M = _model(m, [mat])
    G = np.zeros((m.nb_elements * Q, 9))
    G[:, 0] = 0.2   # W ~ 0.5 (lambda + 2 mu) 0.2^2 ~ 270 > Yd
    S1 = M.material_compute_stress(0, G)
    d1 = M.field("damage").copy()
    assert np.all(d1 > 0) and np.all(d1 <= 1)
    Sp, dp = orc.point_stress(orc.NEOHOOKEAN_DAMAGE, 3, 1e4, 0.3, G[0].reshape(3, 3).T,
                              Sd=100.0, Yd=10.0)
    assert d1[0] == dp
    assert np.array_equal(S1[0].reshape(3, 3).T, Sp)
    # a smaller strain afterwards keeps the damage (history)
    S2 = M.material_compute_stress(0, 0.5 * G)
    assert np.array_equal(M.field("damage"), d1)
    assert np.all(np.abs(S2) <= np.abs(S1) + 1e-9)
