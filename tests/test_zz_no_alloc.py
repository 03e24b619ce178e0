"""akx_energy_noalloc (the allocation-free variant of akx_energy) through the C
ABI against akx_energy on the same device arrays: bit-identical kinetic
energy and external work, on two calls in a row, the second with a smaller
input, in 2D and 3D."""
import ctypes

import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm

pytestmark = pytest.mark.gpu


def _model(mesh, seed):
    g = akx.SolidMechanicsModel(mesh)
    g.add_material(E=210e9, nu=0.3, rho=7800.0)
    g.init_full()
    rng = np.random.default_rng(seed)
    X = mesh.nodes
    blocked = np.zeros_like(X, dtype=bool)
    blocked[np.isclose(X[:, 0], 0.0)] = True
    g.set("blocked_dofs", blocked)
    g.set("velocity", rng.standard_normal(X.shape))
    g.set("external_force", rng.standard_normal(X.shape))
    g.set_time_step(0.5 * g.get_stable_time_step())
    g.solve_step(3, akx.FP64)   # internal force of the last step, moving blocked DOFs
    v = g.get_velocity()
    v[blocked] = 0.3
    g.set("velocity", v)
    return g


def _ptr(g, name):
    p = ctypes.c_void_p()
    g.ctx.check(akx.lib.akx_field_device_pointer(g.ctx.h, akx.FIELDS[name], ctypes.byref(p)))
    return p


def _hip():
    """The HIP runtime libakx.so itself loaded (its path from the process's
    memory map), for the test's own stream and output buffer."""
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libakx.so has not loaded the HIP runtime"
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(hip, g, kind, out, stream):
    rc = akx.lib.akx_energy_noalloc(
        kind, g.dim, g.mesh.nb_nodes, _ptr(g, "velocity"), _ptr(g, "mass"),
        _ptr(g, "external_force"), _ptr(g, "internal_force"), _ptr(g, "blocked_dofs"),
        g.time_step, out, stream)
    assert rc == 0
    assert hip.hipStreamSynchronize(stream) == 0
    v = ctypes.c_double(float("nan"))
    assert hip.hipMemcpy(ctypes.byref(v), out, 8, 2) == 0   # device to host
    return v.value


def test_energy_noalloc_equals_energy():
    meshes = [akm.bar_tri3(300, 60), akm.bar_tri3(30, 6), akm.cube_tet4(12, 0.01, 1),
              akm.cube_tet4(4, 0.01, 2)]
    models = [_model(mesh, k) for k, mesh in enumerate(meshes)]
    hip = _hip()
    out, stream = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), 8) == 0
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # larger, then smaller inputs, 2D then 3D: the scratch array keeps the
        # last call's partial sums, which must not leak into the next result
        for k, g in enumerate(models):
            g.synchronize()
            for kind, name in ((0, "kinetic"), (2, "external work")):
                ref = g.get_energy(name)
                got = _noalloc(hip, g, kind, out, stream)
                assert ref != 0.0
                assert got == ref, (k, name, got, ref)
    finally:
        hip.hipStreamDestroy(stream)
        hip.hipFree(out)
        for g in models:
            g.close()


def test_energy_noalloc_rejects_bad_arguments():
    out = ctypes.c_void_p(1)
    v = ctypes.c_void_p(1)
    assert akx.lib.akx_energy_noalloc(1, 3, 10, v, v, v, v, v, 1.0, out, None) == -5
    assert akx.lib.akx_energy_noalloc(0, 4, 10, v, v, v, v, v, 1.0, out, None) == -1
    assert akx.lib.akx_energy_noalloc(0, 3, -1, v, v, v, v, v, 1.0, out, None) == -1
    assert akx.lib.akx_energy_noalloc(0, 3, 10, v, None, v, v, v, 1.0, out, None) == -1
