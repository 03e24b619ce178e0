"""The reference's energy and dynamics tests (restated in
tests/explicit_drivers.py) on the GPU through the C ABI: kinetic energy
under a uniform velocity, the explicit work balance with a traction and with
a boundary moved at a prescribed velocity, and the analytic plane wave.

FP64: every figure the driver returns within 1e-12 of the oracle's (same
mesh, same steps), and the reference's own tolerance.  Mixed: the
reference's tolerance."""
import os

import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from tests import explicit_drivers as drv

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-12):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("name", ["triangle_3", "quadrangle_4", "tetrahedron_4",
                                  "hexahedron_8"])
def test_kinetic_energy_uniform_velocity(golden_dir, name):
    mesh = akm.read_msh(os.path.join(golden_dir, "patch_%s.msh" % name))
    for precision in (akx.FP64, akx.MIXED):
        ek = drv.kinetic_energy_run(drv.GpuEngine(mesh, precision=precision), mesh)
        assert np.abs(ek - 0.5).max() <= 1e-8, precision


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_work_dynamics(et):
    mesh = drv.bar(et)
    ref = drv.work_dynamics_run(drv.OracleEngine(mesh), mesh)
    fp64 = drv.work_dynamics_run(drv.GpuEngine(mesh, precision=akx.FP64), mesh)
    assert _close(fp64, ref)
    mixed = drv.work_dynamics_run(drv.GpuEngine(mesh, precision=akx.MIXED), mesh)
    for W, dE in mixed:
        assert abs(W - dE) <= 5e-2, (W, dE)
    assert np.abs(mixed - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_work_prescribed_velocity(et):
    mesh = drv.bar(et)
    ref = drv.prescribed_velocity_run(drv.OracleEngine(mesh), mesh)
    fp64 = drv.prescribed_velocity_run(drv.GpuEngine(mesh, precision=akx.FP64), mesh)
    assert _close(fp64, ref)
    W, dE = drv.prescribed_velocity_run(drv.GpuEngine(mesh, precision=akx.MIXED), mesh)
    assert abs(W - dE) <= 1e-2 * abs(dE)


@pytest.mark.parametrize("et", drv.ELEMENT_TYPES)
def test_plane_wave(et):
    mesh = drv.bar(et)
    ref, steps = drv.plane_wave_run(drv.OracleEngine(mesh), mesh)
    fp64, s64 = drv.plane_wave_run(drv.GpuEngine(mesh, precision=akx.FP64), mesh)
    assert s64 == steps and _close(fp64, ref)
    mixed, _ = drv.plane_wave_run(drv.GpuEngine(mesh, precision=akx.MIXED), mesh)
    assert 0 < mixed <= 2e-3
    assert abs(mixed - ref) <= 1e-5 * ref
