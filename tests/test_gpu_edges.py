"""Edge cases of the GPU path: single-element meshes of every element type
through every kernel variant (one partly filled patch, one node block), and
the C ABI's argument and state errors.  Ragged patches (element counts that
are not a multiple of the patch size) and patches too large for the packed
16-bit indices are covered by the c0 bar (4,000 triangles) and the hex8 cube
in test_gpu_parity.py."""
import ctypes

import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from tests import gpu_helpers as gh

pytestmark = pytest.mark.gpu

ONE_ELEMENT = {
    "triangle_3": (akm.TRIANGLE_3, [[0.0, 0.0], [1.0, 0.1], [0.2, 0.9]], [[0, 1, 2]]),
    "quadrangle_4": (akm.QUADRANGLE_4, [[0.0, 0.0], [1.0, 0.0], [1.1, 0.9], [0.0, 1.0]],
                     [[0, 1, 2, 3]]),
    "tetrahedron_4": (akm.TETRAHEDRON_4,
                      [[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.1, 1.0, 0.1], [0.0, 0.2, 1.0]],
                      [[0, 1, 2, 3]]),
    "hexahedron_8": (akm.HEXAHEDRON_8,
                     [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1.1], [0, 1, 1]],
                     [[0, 1, 2, 3, 4, 5, 6, 7]]),
}


def _one_element_pair(name, kind=akx.ELASTIC):
    et, nodes, conn = ONE_ELEMENT[name]
    m = akm.Mesh(et, np.array(nodes, dtype=float), np.array(conn))
    p = gh.Pair(m, [dict(kind=kind, E=1e7, nu=0.3, rho=1000.0)])
    rng = np.random.default_rng(5)
    p.set_nodal("velocity", rng.normal(scale=0.1, size=m.nodes.shape))
    mask = np.zeros_like(m.nodes, dtype=bool)
    mask[0] = True                      # one node held: a forced, non-rigid motion
    p.block(mask)
    p.set_time_step(0.5 * p.cpu.stable_time_step())
    return p


@pytest.mark.parametrize("name", sorted(ONE_ELEMENT))
def test_single_element_fp64_every_layout(name):
    """FP64 on one element == oracle bit for bit, whatever the launch shape
    (layouts 2-5 fall back to node-ordered slots on the FP64 path, 6 is the
    fused node-block step)."""
    for layout in range(7):
        for recompute in (0, 1):
            p = _one_element_pair(name)
            p.gpu.set_launch(64, 64, 1, recompute, layout, 0)
            p.step(40, akx.FP64)
            assert np.array_equal(p.gpu.get_displacement(), p.cpu.nodal("u")), (layout, recompute)
            assert np.array_equal(p.gpu.get_velocity(), p.cpu.nodal("v")), (layout, recompute)


@pytest.mark.parametrize("name", sorted(ONE_ELEMENT))
@pytest.mark.parametrize("kind", [akx.ELASTIC, akx.NEOHOOKEAN])
def test_single_element_mixed_every_layout(name, kind):
    """Mixed path on one element (a patch of one element, a node block of a
    few nodes): <= 1e-5 from the oracle after 40 steps for every layout
    (layout 6 on triangles: from the per-evaluation records and through
    element ids)."""
    for layout in range(7):
        for recompute in ((0, 1) if layout < 2 or layout == 6 else (1,)):
            p = _one_element_pair(name, kind)
            p.gpu.set_launch(256, 256, 1, recompute, layout, 0)
            p.step(40, akx.MIXED)
            err = gh.rel(p.gpu.get_displacement(), p.cpu.nodal("u"))
            assert err <= 1e-5, (layout, recompute, err)


FINITE_STRAIN_MESHES = {
    "triangle_3": lambda: akm.rectangle_tri3(4, 4, 1.0, 1.0),
    "quadrangle_4": lambda: akm.rectangle_quad4(4, 4, 1.0, 1.0),
    "tetrahedron_4": lambda: akm.jitter_interior(akm.box_tet4(3, 3, 3), 0.05, 3),
    "hexahedron_8": lambda: akm.jitter_interior(akm.box_hex8(3, 3, 3), 0.05, 3),
}


@pytest.mark.parametrize("name", sorted(FINITE_STRAIN_MESHES))
@pytest.mark.parametrize("stretch", [0.02, 0.15])
def test_mixed_neohookean_forces_at_finite_strain(name, stretch):
    """Mixed-path neo-Hookean internal forces under a rotated finite stretch
    (30 degrees, principal stretches up to 1 + `stretch`, plus a random
    non-uniform part) against the oracle, <= 1e-5, for the table and rebuilt-
    geometry kernels and the patch layouts.  At stretch 0.15, |J - 1| > 1/16
    takes the FP32 law's log1pf branch, at 0.02 its series."""
    mesh = FINITE_STRAIN_MESHES[name]()
    X = mesh.nodes
    d = X.shape[1]
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    if d == 2:
        R = np.array([[c, -s], [s, c]])
        U = np.diag([1 + stretch, 1 - stretch / 2])
    else:
        k = np.ones(3) / np.sqrt(3.0)
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * K @ K
        U = np.diag([1 + stretch, 1 - stretch / 2, 1 + stretch / 3])
    rng = np.random.default_rng(11)
    u = X @ (R @ U - np.eye(d)).T + rng.normal(scale=0.02 * stretch, size=X.shape)
    p = gh.Pair(mesh, [dict(kind=akx.NEOHOOKEAN, E=1e7, nu=0.3, rho=1000.0)])
    p.set_nodal("displacement", u)
    p.cpu.assemble_internal_forces()
    ref = p.cpu.nodal("f_int").copy()
    for layout, recompute in ((0, 0), (0, 1), (1, 1), (3, 1), (4, 1), (5, 1)):
        p.gpu.set_launch(256, 256, 1, recompute, layout, 0)
        p.gpu.assemble_internal_forces(akx.MIXED)
        err = gh.rel(p.gpu.get_internal_force(), ref)
        assert err <= 1e-5, (layout, recompute, err)


@pytest.mark.parametrize("keep_numbering", [True, False])
def test_node_without_elements(keep_numbering):
    """A node that no element uses (here the last node, and kept last when the
    library keeps the caller's numbering) has an empty force range: FP64 ==
    oracle bit for bit and mixed <= 1e-5 for every layout, the blocked orphan
    stays at rest."""
    base = akm.cube_tet4(3, 0.02, 5)
    mesh = akm.Mesh(base.elem_type, np.vstack([base.nodes, [[2.0, 2.0, 2.0]]]), base.conn)
    X = mesh.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 2], 0.0)] = True
    mask[-1] = True
    v = np.zeros_like(X)
    v[:, 0] = np.sin(3 * X[:, 1]) * X[:, 2]
    v[-1] = 0.0
    for precision in (akx.FP64, akx.MIXED):
        for layout in range(7):
            p = gh.Pair(mesh, [dict(E=210e9, nu=0.3, rho=7800.0)], keep_numbering=keep_numbering)
            p.block(mask)
            p.set_nodal("velocity", v)
            p.set_time_step(0.8 * p.cpu.stable_time_step())
            p.gpu.set_launch(256, 256, 1, 1, layout, 0)
            p.step(30, precision)
            u = p.gpu.get_displacement()
            assert np.all(u[-1] == 0.0), (precision, layout)
            if precision == akx.FP64:
                assert np.array_equal(u, p.cpu.nodal("u")), layout
            else:
                assert gh.rel(u, p.cpu.nodal("u")) <= 1e-5, layout


def _check(rc, expected):
    assert rc == expected, (rc, expected)


def test_abi_rejects_bad_arguments_and_states():
    lib = akx.lib
    h = ctypes.c_void_p()
    _check(lib.akx_ctx_create(0, 0, ctypes.byref(h)), 0)
    try:
        nodes = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
        bad = np.array([[0, 1, 3]], dtype=np.uint32)       # node 3 does not exist
        _check(lib.akx_mesh_set(h, akm.TRIANGLE_3, 2, 3, nodes.ctypes.data_as(ctypes.c_void_p),
                                1, bad.ctypes.data_as(ctypes.c_void_p)), -1)
        _check(lib.akx_mesh_set(h, akm.TRIANGLE_3, 3, 3, nodes.ctypes.data_as(ctypes.c_void_p),
                                1, bad.ctypes.data_as(ctypes.c_void_p)), -1)   # wrong dimension
        conn = np.array([[0, 1, 2]], dtype=np.uint32)
        _check(lib.akx_mesh_set(h, akm.TRIANGLE_3, 2, 3, nodes.ctypes.data_as(ctypes.c_void_p),
                                1, conn.ctypes.data_as(ctypes.c_void_p)), 0)
        # stepping and asynchronous downloads need akx_init
        _check(lib.akx_solve_step(h, 1, akx.FP64), -2)
        out = np.zeros((3, 2))
        t = ctypes.c_uint64()
        _check(lib.akx_field_download_async(h, 0, out.ctypes.data_as(ctypes.c_void_p),
                                            out.nbytes, ctypes.byref(t)), -2)
        mid = ctypes.c_int()
        _check(lib.akx_material_add(h, akx.ELASTIC, 1e7, 0.3, 1000.0, 0, ctypes.byref(mid)), 0)
        _check(lib.akx_init(h), 0)
        _check(lib.akx_solve_step(h, 1, akx.FP64), -2)        # no time step yet
        _check(lib.akx_set_time_step(h, -1.0), -1)
        _check(lib.akx_set_time_step(h, 1e-6), 0)
        # sizes must match exactly; layouts are 0..6; tickets must exist
        _check(lib.akx_field_upload(h, 0, out.ctypes.data_as(ctypes.c_void_p), out.nbytes - 8), -1)
        _check(lib.akx_field_download(h, 0, out.ctypes.data_as(ctypes.c_void_p), out.nbytes + 8), -1)
        _check(lib.akx_set_launch(h, 0, 0, -1, -1, 7, -1), -1)
        _check(lib.akx_download_wait(h, 12345), -1)
        done = ctypes.c_int()
        _check(lib.akx_download_query(h, 0, ctypes.byref(done)), -1)
        # a static table is not available asynchronously
        jac = np.zeros(1)
        _check(lib.akx_field_download_async(h, 13, jac.ctypes.data_as(ctypes.c_void_p), 8,
                                            ctypes.byref(t)), -1)
        assert b"asynchronously" in lib.akx_last_error(h)
        _check(lib.akx_solve_step(h, 1, 7), -1)             # unknown precision
    finally:
        _check(lib.akx_ctx_destroy(h), 0)
