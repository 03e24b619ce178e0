"""The akantu:: C++ API (libakantu_amd.so) and its example drivers.

The drivers follow the reference's own C++ programs for this path:
  * facade/examples/patch_test_explicit.cc:
      test/test_model/patch_tests/patch_test_linear_elastic_explicit.cc:33-68;
  * facade/examples/explicit_dynamic.cc:
      examples/explicit/explicit_dynamic.cc:38-106 on the bar of
      examples/explicit/bar.geo (10 x 1 x 1, x from -5) with
      examples/explicit/material.dat, built here as a structured tet4 box;
  * facade/examples/fe_engine_operators.cc: the FEEngine checks of
      test/test_fe_engine/test_gradient.cc:46-102 and test_integrate.cc:41-72,
      the lumped mass, and SolidMechanicsModel::assembleInternalForces
      restated on the Material API (computeAllStresses, assembleInternalForces
      material by material).
The GPU tests run them as child processes and check what they write against
the FP64 oracle on the same mesh and initial state.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from akantu_amd import mesh as akm
from tests import patch_helpers as ph

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "akantu_amd")
BIN = os.path.join(PKG, "facade", "bin")
LIB = os.path.join(PKG, "libakantu_amd.so")

STEEL_PLANE_STRAIN = """material elastic [
  name = steel
  rho = 7800
  E = 2.1e11
  nu = 0.3
  Plane_Stress = false
]
"""

# examples/explicit/material.dat
BAR_STEEL = """material elastic [
  name = steel
  rho = 7800   # density
  E   = 2.1e11 # Young's modulus
  nu  = 0.0    # Poisson's ratio
]
"""


def _built():
    return os.path.exists(LIB) and all(
        os.path.exists(os.path.join(BIN, b))
        for b in ("explicit_dynamic", "patch_test_explicit", "fe_engine_operators"))


needs_build = pytest.mark.skipif(not _built(), reason="C++ API not built (akantu_amd/build.py)")


@needs_build
def test_facade_library_loads():
    ctypes.CDLL(LIB)


@needs_build
@pytest.mark.parametrize("prog", ["explicit_dynamic", "patch_test_explicit", "fe_engine_operators"])
def test_driver_usage(prog):
    r = subprocess.run([os.path.join(BIN, prog)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _run(cmd, cwd):
    print("running:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=cwd)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    return r


@pytest.mark.gpu
@needs_build
@pytest.mark.parametrize("name", ph.PATCH_TYPES)
def test_patch_test_driver(name, tmp_path):
    mat = tmp_path / "material.dat"
    mat.write_text(STEEL_PLANE_STRAIN)
    msh = os.path.join(ph.GOLDEN, "patch_%s.msh" % name)
    dim = akm.read_msh(msh).dim
    r = _run([os.path.join(BIN, "patch_test_explicit"), msh, str(mat), str(dim)], tmp_path)
    assert r.returncode == 0
    assert r.stdout.count("PASS") == 5 and "OK" in r.stdout


def _bar():
    return akm.box_tet4(25, 3, 3, lx=10.0, ly=1.0, lz=1.0, origin=(-5.0, 0.0, 0.0))


@pytest.mark.gpu
@needs_build
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_explicit_dynamic_driver_matches_oracle(precision, tmp_path):
    from oracle import oracle as orc

    m = _bar()
    msh = tmp_path / "bar.msh"
    akm.write_msh(m, str(msh))
    (tmp_path / "material.dat").write_text(BAR_STEEL)
    steps = 300
    r = _run([os.path.join(BIN, "explicit_dynamic"), str(msh), str(tmp_path / "material.dat"),
              "3", str(steps), precision, str(tmp_path) + "/", "100"], tmp_path)
    assert r.returncode == 0
    # ParaView dumps: the initial state and every 100 steps
    dumps = sorted(p.name for p in (tmp_path / "paraview").glob("explicit_dynamic_*.vtu"))
    assert len(dumps) == 1 + steps // 100
    assert (tmp_path / "paraview" / "explicit_dynamic.pvd").exists()
    u0 = np.loadtxt(tmp_path / "u0.txt")
    u = np.loadtxt(tmp_path / "u.txt")
    # the last dump (written by the background writer) holds the final state
    import base64
    import xml.etree.ElementTree as ET
    root = ET.parse(tmp_path / "paraview" / dumps[-1]).getroot()
    assert root.get("header_type") == "UInt64"
    piece = root.find("UnstructuredGrid/Piece")
    kinds = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}
    arrays = {}
    for da in piece.iter("DataArray"):
        assert da.get("format") == "binary"
        raw = base64.b64decode(da.text.strip())
        assert int(np.frombuffer(raw[:8], dtype="<u8")[0]) == len(raw) - 8
        arrays[da.get("Name")] = np.frombuffer(raw[8:], dtype=kinds[da.get("type")])
    assert np.array_equal(arrays["displacement"].reshape(-1, 3), u)
    assert arrays["stress"].size == 9 * m.nb_elements
    assert np.array_equal(arrays["connectivity"].reshape(m.conn.shape), m.conn)
    assert np.array_equal(arrays["offsets"], m.conn.shape[1] * np.arange(1, m.nb_elements + 1))
    dt = float((tmp_path / "dt.txt").read_text())
    en = np.loadtxt(tmp_path / "energy.csv", delimiter=",", skiprows=1, usecols=(0, 2, 3))
    assert en.shape == (steps, 3)

    cpu = orc.OracleModel(m, [dict(E=2.1e11, nu=0.0, rho=7800.0)])
    assert dt == 0.8 * cpu.stable_time_step()
    cpu.set_time_step(dt)
    cpu.nodal("u")[:] = u0
    cpu.solve_step(steps)
    ref = cpu.nodal("u")
    ek, ep = cpu.energy("kinetic"), cpu.energy("potential")
    if precision == "fp64":
        # same operations in the same order: bit-identical displacements
        assert np.array_equal(u, ref)
        tol = 1e-12
    else:
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-5
        tol = 1e-5
    assert abs(en[-1, 1] - ep) <= tol * (ep + ek)
    assert abs(en[-1, 2] - ek) <= tol * (ep + ek)
    # the energy of the free bar is conserved by the central-difference scheme
    tot = en[:, 1] + en[:, 2]
    assert np.abs(tot - tot[0]).max() < 1e-2 * tot[0]


TWO_MATERIALS = """material elastic [
  name = steel
  rho = 7800
  E = 2.1e11
  nu = 0.3
]
material neohookean [
  name = rubber
  rho = 1000
  E = 1e7
  nu = 0.45
]
"""


@pytest.mark.gpu
@needs_build
@pytest.mark.parametrize("name", ["triangle_3", "quadrangle_4", "tetrahedron_4", "hexahedron_8"])
def test_fe_engine_operators_driver(name, tmp_path):
    """The FEEngine / Material driver on the reference's FEEngine test meshes,
    two materials (elastic, neo-Hookean) on alternating elements; what it
    writes against the oracle."""
    from oracle import oracle as orc

    msh = os.path.join(ph.GOLDEN, "fe_%s.msh" % name)
    m = akm.read_msh(msh)
    (tmp_path / "material.dat").write_text(TWO_MATERIALS)
    r = _run([os.path.join(BIN, "fe_engine_operators"), msh, str(tmp_path / "material.dat"),
              str(m.dim), str(tmp_path / "out")], tmp_path)
    assert r.returncode == 0 and "OK" in r.stdout
    d, Q, E = m.dim, m.nb_qp, m.nb_elements
    em = (np.arange(E) % 3 == 1).astype(np.int32)
    mats = [dict(kind=orc.ELASTIC, E=2.1e11, nu=0.3, rho=7800.0),
            dict(kind=orc.NEOHOOKEAN, E=1e7, nu=0.45, rho=1000.0)]
    cpu = orc.OracleModel(m, mats, em)
    u = np.fromfile(str(tmp_path / "out_u.bin")).reshape(-1, d)
    assert np.abs(u).max() > 0
    grad = np.fromfile(str(tmp_path / "out_gradient.bin")).reshape(E * Q, d * d)
    assert np.array_equal(grad, cpu.fe_gradient(u, d))
    lumped = np.fromfile(str(tmp_path / "out_lumped.bin"))
    assert np.array_equal(lumped, cpu.nodal("mass")[:, 0])
    cpu.nodal("u")[:] = u
    cpu.assemble_internal_forces()
    f = np.fromfile(str(tmp_path / "out_fint.bin")).reshape(-1, d)
    ref = cpu.nodal("f_int")
    assert np.abs(f - ref).max() <= 1e-12 * np.abs(ref).max()
