"""ParaView output (.vtu / .pvd) written by akantu_amd/dumper.py."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from akantu_amd import dumper
from akantu_amd import mesh as akm


def _arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    for da in piece.iter("DataArray"):
        vals = np.array([float(v) for v in da.text.split()])
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    return piece, out


@pytest.mark.parametrize("make", [lambda: akm.rectangle_tri3(3, 2, 1.0, 0.5),
                                  lambda: akm.box_hex8(2, 2, 1)])
def test_vtu_round_trip(tmp_path, make):
    m = make()
    rng = np.random.default_rng(0)
    u = rng.normal(size=m.nodes.shape)
    s = rng.normal(size=(m.nb_elements, m.dim * m.dim))
    path = tmp_path / "out" / "a_0000.vtu"
    dumper.write_vtu(str(path), m, {"displacement": u}, {"stress": s})
    piece, arr = _arrays(path)
    assert int(piece.get("NumberOfPoints")) == m.nb_nodes
    assert int(piece.get("NumberOfCells")) == m.nb_elements
    assert np.array_equal(arr["connectivity"].astype(np.int64), m.conn.reshape(-1).astype(np.int64))
    assert np.allclose(arr["positions"][:, :m.dim], m.nodes, rtol=0, atol=0)
    got = arr["displacement"][:, :m.dim]
    assert np.array_equal(got, u)           # repr() keeps every bit
    assert np.array_equal(arr["stress"], s)
    assert set(arr["types"]) == {dumper.VTK_CELL[m.elem_type]}


def test_pvd_lists_the_dumps(tmp_path):
    dumper.write_pvd(str(tmp_path / "b.pvd"), [(0.0, "b_0000.vtu"), (1e-3, "b_0001.vtu")])
    ds = ET.parse(tmp_path / "b.pvd").getroot().findall("Collection/DataSet")
    assert [d.get("file") for d in ds] == ["b_0000.vtu", "b_0001.vtu"]
    assert float(ds[1].get("timestep")) == 1e-3


@pytest.mark.gpu
def test_model_dump_writes_the_device_state(tmp_path):
    from akantu_amd import akx
    m = akm.bar_tri3(20, 4)
    model = akx.SolidMechanicsModel(m)
    model.add_material(E=210e9, nu=0.3, rho=7800.0)
    model.init_full()
    v0 = np.zeros_like(m.nodes)
    v0[:, 0] = 0.01 * np.sin(np.pi * m.nodes[:, 0] / 2)
    model.set("velocity", v0)
    model.set_time_step(0.8 * model.get_stable_time_step())
    model.set_base_name("bar", str(tmp_path / "paraview"))
    for f in ("displacement", "velocity", "stress"):
        model.add_dump_field(f)
    model.solve_step(20, akx.FP64)
    path = model.dump()
    piece, arr = _arrays(path)
    assert np.array_equal(arr["displacement"][:, :2], model.get_displacement())
    assert np.array_equal(arr["velocity"][:, :2], model.get_velocity())
    assert np.array_equal(arr["stress"], model.get("stress"))
    model.solve_step(20, akx.FP64)
    model.dump()
    ds = ET.parse(tmp_path / "paraview" / "bar.pvd").getroot().findall("Collection/DataSet")
    assert len(ds) == 2
