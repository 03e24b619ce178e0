"""Pins of the oracle's stable time step (element inradius) and energies.

* tetrahedron_4 2·r_in (element_class_tetrahedron_4_inline_impl.hh:128-134,
  aka_math_tmpl.hh:650-679): a/sqrt(6) for the regular tetrahedron of edge
  a; h·(sqrt(2) - 1) for every Kuhn tetrahedron of an h-cube; 6V / (sum of
  face areas) with V and the areas from cross products on random
  tetrahedra (the oracle uses a determinant expansion and Heron's formula);
* quadrangle_4 / hexahedron_8: the shortest edge
  (element_class_quadrangle_4_inline_impl.hh:131-154,
  element_class_hexahedron_8_inline_impl.hh:218-246), with the edges found
  as the node pairs whose natural coordinates differ in one coordinate, and
  each edge made the unique shortest in turn;
* kinetic energy, external work and the explicit dynamics through the
  reference's own tests, restated in tests/explicit_drivers.py;
* each check is shown to catch a planted mistake: copies of the oracle
  built with one term altered fail it (the mistake is planted on the CPU,
  in a temporary copy of the oracle's source).
"""
import os
import subprocess

import numpy as np
import pytest

from akantu_amd import mesh as akm
from oracle import oracle as orc
from tests import explicit_drivers as drv

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle", "akx_oracle.cc")

# natural coordinates of the nodes, in the reference's node order
QUAD_XI = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
HEX_XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def _edges(xi):
    n = xi.shape[0]
    return [(a, b) for a in range(n) for b in range(a + 1, n)
            if int(np.sum(xi[a] != xi[b])) == 1]


def element_size(et, X, library=None):
    """The oracle's stable time step of one element with c_p = 1 (E = rho = 1,
    nu = 0): the element's characteristic size."""
    n = {akm.TRIANGLE_3: 3, akm.QUADRANGLE_4: 4, akm.TETRAHEDRON_4: 4,
         akm.HEXAHEDRON_8: 8}[et]
    m = akm.Mesh(et, np.asarray(X, dtype=float), np.arange(n, dtype=np.uint32)[None, :])
    return orc.OracleModel(m, [drv.UNIT], library=library).stable_time_step()


def _tet_closed_form(P):
    V = abs(np.dot(P[1] - P[0], np.cross(P[2] - P[0], P[3] - P[0]))) / 6.0
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    A = sum(0.5 * np.linalg.norm(np.cross(P[j] - P[i], P[k] - P[i])) for i, j, k in faces)
    return 6.0 * V / A


def _positive_tet(rng):
    while True:
        P = rng.uniform(-1, 1, size=(4, 3))
        vol = np.dot(P[1] - P[0], np.cross(P[2] - P[0], P[3] - P[0])) / 6.0
        if vol < -0.05:   # the reference's tetrahedron_4 orientation has det J < 0 here
            P[[2, 3]] = P[[3, 2]]
        if abs(vol) > 0.05:
            return P


def check_tet_inradius(library=None):
    a = 0.7
    reg = a * np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0],
                        [0.5, np.sqrt(3) / 6, np.sqrt(2.0 / 3.0)]])
    assert element_size(akm.TETRAHEDRON_4, reg, library) == pytest.approx(a / np.sqrt(6),
                                                                          rel=1e-14)
    h = 0.25
    cube = akm.box_tet4(1, 1, 1, h, h, h)
    for tet in cube.conn:
        size = element_size(akm.TETRAHEDRON_4, cube.nodes[tet], library)
        assert size == pytest.approx(h * (np.sqrt(2) - 1), rel=1e-14)
    rng = np.random.default_rng(5)
    for _ in range(20):
        P = _positive_tet(rng)
        assert element_size(akm.TETRAHEDRON_4, P, library) == pytest.approx(
            _tet_closed_form(P), rel=1e-12)


def _shortest_edge_cases(xi):
    """For each edge (a, b) of the reference element: the element scaled to
    edge 1 with node b moved halfway towards node a, so that this edge is
    the unique shortest (0.5; the others are 1 or longer)."""
    base = 0.5 * (xi + 1)
    for a, b in _edges(xi):
        X = base.copy()
        X[b] = 0.5 * (X[a] + X[b])
        yield (a, b), X


def check_shortest_edge(et, library=None):
    xi = QUAD_XI if et == akm.QUADRANGLE_4 else HEX_XI
    nedges = 0
    for edge, X in _shortest_edge_cases(xi):
        assert element_size(et, X, library) == pytest.approx(0.5, rel=1e-15), edge
        nedges += 1
    assert nedges == (4 if et == akm.QUADRANGLE_4 else 12)
    # a distorted element: the minimum over its edges
    rng = np.random.default_rng(9)
    X = 0.5 * (xi + 1) + rng.uniform(-0.15, 0.15, size=xi.shape)
    expect = min(np.linalg.norm(X[a] - X[b]) for a, b in _edges(xi))
    assert element_size(et, X, library) == pytest.approx(expect, rel=1e-15)


def check_work_dynamics(et, library=None):
    mesh = drv.bar(et)
    out = drv.work_dynamics_run(drv.OracleEngine(mesh, library=library), mesh)
    for W, dE in out:
        assert abs(W) > 0.5
        assert abs(W - dE) <= 5e-2, (W, dE)


def check_prescribed_velocity(et, library=None):
    mesh = drv.bar(et)
    W, dE = drv.prescribed_velocity_run(drv.OracleEngine(mesh, library=library), mesh)
    assert dE > 0
    assert abs(W - dE) <= 1e-2 * abs(dE), (W, dE)


def test_tet_inradius():
    check_tet_inradius()


@pytest.mark.parametrize("et", [akm.QUADRANGLE_4, akm.HEXAHEDRON_8])
def test_shortest_edge(et):
    check_shortest_edge(et)


def test_triangle_inradius_closed_form():
    """2·r_in of a right triangle with legs p, q: 2·pq / (p + q + hypot)."""
    p, q = 0.3, 0.4
    X = np.array([[0, 0], [p, 0], [0, q]])
    assert element_size(akm.TRIANGLE_3, X) == pytest.approx(2 * p * q / (p + q + 0.5),
                                                            rel=1e-14)


@pytest.mark.parametrize("name", ["triangle_3", "quadrangle_4", "tetrahedron_4",
                                  "hexahedron_8"])
def test_kinetic_energy_uniform_velocity(golden_dir, name):
    mesh = akm.read_msh(os.path.join(golden_dir, "patch_%s.msh" % name))
    eng = drv.OracleEngine(mesh)
    assert eng.get("v").shape == mesh.nodes.shape
    ek = drv.kinetic_energy_run(eng, mesh)
    assert np.abs(ek - 0.5).max() <= 1e-8


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_work_dynamics(et):
    check_work_dynamics(et)


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_work_prescribed_velocity(et):
    check_prescribed_velocity(et)


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_plane_wave(et):
    mesh = drv.bar(et)
    worst, steps = drv.plane_wave_run(drv.OracleEngine(mesh), mesh)
    assert steps >= 100
    assert 0 < worst <= 2e-3


# ---------------------------------------------------------------------------
# planted mistakes: each must make its check fail
MUTANTS = {
    # a face of the tetrahedron with a wrong edge
    "tet_face": ("s2 = (l12 + l24 + l14) * 0.5; s2 = std::sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l14));",
                 "s2 = (l12 + l24 + l13) * 0.5; s2 = std::sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l13));",
                 lambda L: check_tet_inradius(L)),
    "tet_edge": ("double l23 = dist3d(c2, c3), l24 = dist3d(c2, c4), l34 = dist3d(c3, c4);",
                 "double l23 = dist3d(c2, c3), l24 = dist3d(c1, c4), l34 = dist3d(c3, c4);",
                 lambda L: check_tet_inradius(L)),
    "quad_edge": ("double c = dist2d(x + 4, x + 6), d = dist2d(x + 6, x + 0);",
                  "double c = dist2d(x + 4, x + 6), d = dist2d(x + 6, x + 2);",
                  lambda L: check_shortest_edge(akm.QUADRANGLE_4, L)),
    "hex_edge": ("dist3d(P(2), P(6)), dist3d(P(3), P(7)), dist3d(P(4), P(5)),",
                 "dist3d(P(2), P(6)), dist3d(P(3), P(6)), dist3d(P(4), P(5)),",
                 lambda L: check_shortest_edge(akm.HEXAHEDRON_8, L)),
    "work_free_term": ("else work += M->fext[k] * M->v[k];",
                       "else work += M->fext[k] * M->u[k];",
                       lambda L: check_work_dynamics(akm.TRIANGLE_3, L)),
    "work_blocked_sign": ("if (M->blocked[k]) work -= M->fint[k] * M->v[k];",
                          "if (M->blocked[k]) work += M->fint[k] * M->v[k];",
                          lambda L: check_prescribed_velocity(akm.TRIANGLE_3, L)),
    "work_time_step": ("return work * M->dt;", "return work;",
                       lambda L: check_work_dynamics(akm.QUADRANGLE_4, L)),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(tmp_path, name):
    before, after, check = MUTANTS[name]
    src = open(ORACLE_SRC).read()
    assert src.count(before) == 1, "the term to alter is no longer in the oracle"
    mutated = tmp_path / "mutant.cc"
    mutated.write_text(src.replace(before, after))
    so = tmp_path / "libmutant.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread",
                    "-shared", "-o", str(so), str(mutated)], check=True)
    with pytest.raises(AssertionError):
        check(orc.load(str(so)))
