"""ParaView output (.vtu / .pvd) written by akantu_amd/dumper.py."""
import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from akantu_amd import dumper
from akantu_amd import mesh as akm


_DT = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}


def _arrays(path):
    """Every DataArray of a .vtu, decoded from ASCII or from VTK's inline
    binary form (base64 of a UInt64 byte count followed by the raw data)."""
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    for da in piece.iter("DataArray"):
        if da.get("format") == "binary":
            assert root.get("header_type") == "UInt64"
            raw = base64.b64decode(da.text)
            nbytes = int(np.frombuffer(raw[:8], dtype="<u8")[0])
            assert len(raw) == 8 + nbytes
            vals = np.frombuffer(raw[8:], dtype=_DT[da.get("type")]).astype(np.float64)
        else:
            vals = np.array([float(v) for v in da.text.split()])
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    return piece, out


@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("make", [lambda: akm.rectangle_tri3(3, 2, 1.0, 0.5),
                                  lambda: akm.box_hex8(2, 2, 1)])
def test_vtu_round_trip(tmp_path, make, binary):
    m = make()
    rng = np.random.default_rng(0)
    u = rng.normal(size=m.nodes.shape)
    s = rng.normal(size=(m.nb_elements, m.dim * m.dim))
    path = tmp_path / "out" / "a_0000.vtu"
    dumper.write_vtu(str(path), m, {"displacement": u}, {"stress": s}, binary=binary)
    piece, arr = _arrays(path)
    assert int(piece.get("NumberOfPoints")) == m.nb_nodes
    assert int(piece.get("NumberOfCells")) == m.nb_elements
    assert np.array_equal(arr["connectivity"].astype(np.int64), m.conn.reshape(-1).astype(np.int64))
    assert np.allclose(arr["positions"][:, :m.dim], m.nodes, rtol=0, atol=0)
    got = arr["displacement"][:, :m.dim]
    assert np.array_equal(got, u)           # repr() keeps every bit
    assert np.array_equal(arr["stress"], s)
    assert set(arr["types"]) == {dumper.VTK_CELL[m.elem_type]}
    assert np.array_equal(arr["offsets"], m.conn.shape[1] * np.arange(1, m.nb_elements + 1))


def test_binary_vtu_reuses_the_mesh_encoding(tmp_path):
    """Two dumps of one mesh written with the cached mesh encoding are the
    files written without it."""
    m = akm.box_tet4(3, 2, 2)
    u = np.random.default_rng(1).normal(size=m.nodes.shape)
    cache = dumper.mesh_arrays(m, True)
    for k in range(2):
        dumper.write_vtu(str(tmp_path / ("c%d.vtu" % k)), m, {"displacement": u * k},
                         mesh_text=cache)
        dumper.write_vtu(str(tmp_path / ("n%d.vtu" % k)), m, {"displacement": u * k})
        assert (tmp_path / ("c%d.vtu" % k)).read_bytes() == (tmp_path / ("n%d.vtu" % k)).read_bytes()


def test_pvd_lists_the_dumps(tmp_path):
    dumper.write_pvd(str(tmp_path / "b.pvd"), [(0.0, "b_0000.vtu"), (1e-3, "b_0001.vtu")])
    ds = ET.parse(tmp_path / "b.pvd").getroot().findall("Collection/DataSet")
    assert [d.get("file") for d in ds] == ["b_0000.vtu", "b_0001.vtu"]
    assert float(ds[1].get("timestep")) == 1e-3


class _StubBuffer:
    """Host stand-in for akx.PinnedBuffer (no device needed)."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.raw = np.zeros(nbytes, dtype=np.uint8)

    def array(self, dtype, shape):
        return self.raw[:int(np.prod(shape)) * np.dtype(dtype).itemsize].view(dtype).reshape(shape)

    def close(self):
        pass


class _StubModel:
    """Answers the dumper's calls from numpy arrays; an asynchronous download
    completes only when wait() is called, with the values of the moment the
    download was requested (a snapshot), like akx_field_download_async."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.time_step = 1e-3
        self.state = {"displacement": np.zeros(mesh.nodes.shape),
                      "blocked_dofs": np.zeros(mesh.nodes.shape, dtype=bool),
                      "stress": np.zeros((mesh.nb_elements * mesh.nb_qp, mesh.dim ** 2))}
        self.pending = {}
        self.waited = []

    def field_layout(self, name):
        a = self.state[name]
        return a.shape, (np.uint8 if a.dtype == bool else np.float64)

    def get(self, name):
        return self.state[name].copy()

    def get_async(self, name, buf):
        ticket = len(self.pending) + 1
        self.pending[ticket] = (buf, name, self.state[name].copy())
        return ticket

    def wait(self, ticket):
        buf, name, snap = self.pending[ticket]
        shape, dtype = self.field_layout(name)
        buf.array(dtype, shape)[...] = snap.astype(dtype)
        self.waited.append(ticket)

    def step(self, k):
        self.state["displacement"] += k
        self.state["stress"] -= k
        self.state["blocked_dofs"][: k % 3] = True


def test_async_dumper_writes_snapshots_in_order(tmp_path, monkeypatch):
    """The writer thread waits for every download, writes each file from the
    state of its own dump() call although the model moved on, keeps the .pvd
    in dump order and reuses the page-locked buffers."""
    from akantu_amd import akx
    monkeypatch.setattr(akx, "PinnedBuffer", _StubBuffer, raising=False)
    m = akm.rectangle_tri3(3, 2, 1.0, 0.5)
    model = _StubModel(m)
    d = dumper.Dumper(model, "s", str(tmp_path), asynchronous=True, depth=2)
    for f in ("displacement", "blocked_dofs", "stress"):
        d.add_dump_field(f)
    expected = []
    for k in range(5):
        expected.append({f: model.get(f) for f in ("displacement", "stress")})
        d.dump()
        model.step(k + 1)
    d.flush()
    assert sorted(model.waited) == list(range(1, 16))
    for k in range(5):
        _, arr = _arrays(tmp_path / ("s_%04d.vtu" % k))
        assert np.array_equal(arr["displacement"][:, :2], expected[k]["displacement"])
        assert np.array_equal(arr["stress"], expected[k]["stress"])
    ds = ET.parse(tmp_path / "s.pvd").getroot().findall("Collection/DataSet")
    assert [x.get("file") for x in ds] == ["s_%04d.vtu" % k for k in range(5)]
    # at most `depth` dumps in flight: never more than depth + 1 buffers per field
    assert all(len(v) <= 3 for v in d._free.values())
    d.close()


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
    model.flush_dumps()
    piece, arr = _arrays(path)
    assert np.array_equal(arr["displacement"][:, :2], model.get_displacement())
    assert np.array_equal(arr["velocity"][:, :2], model.get_velocity())
    assert np.array_equal(arr["stress"], model.get("stress"))
    # what the files hold is the reference's state: the oracle after the same
    # 20 steps (its stresses are those of the last step's assembly, on the
    # displacement the corrector leaves unchanged)
    from oracle import oracle as orc
    cpu = orc.OracleModel(m, [dict(E=210e9, nu=0.3, rho=7800.0)])
    cpu.nodal("v")[:] = v0
    cpu.set_time_step(model.time_step)
    cpu.solve_step(20)
    assert np.array_equal(arr["displacement"][:, :2], cpu.nodal("u"))
    assert np.array_equal(arr["velocity"][:, :2], cpu.nodal("v"))
    assert np.array_equal(arr["stress"], cpu.field("stress").reshape(m.nb_elements, 4))
    model.solve_step(20, akx.FP64)
    model.dump()
    model.flush_dumps()
    ds = ET.parse(tmp_path / "paraview" / "bar.pvd").getroot().findall("Collection/DataSet")
    assert len(ds) == 2
    model.close()


@pytest.mark.gpu
def test_hex_damage_dump_equals_oracle(tmp_path):
    """Hexahedra with the neo-Hookean + damage law, 30 FP64 steps: the dumped
    nodal fields equal the oracle's, the per-element stress and damage (the
    average of the 8 points) are within 1e-12 (the law's log)."""
    from akantu_amd import akx
    from oracle import oracle as orc
    model = _hex_damage_model()
    m = model.mesh
    rng = np.random.default_rng(1)
    v0 = rng.normal(scale=0.5, size=m.nodes.shape)
    blocked = np.zeros(m.nodes.shape, dtype=bool)
    blocked[m.nodes[:, 2] == 0.0] = True
    model.set_base_name("hex", str(tmp_path / "paraview"))
    for f in ("displacement", "velocity", "stress", "damage"):
        model.add_dump_field(f)
    model.solve_step(30, akx.FP64)
    path = model.dump()
    model.flush_dumps()
    _, arr = _arrays(path)
    cpu = orc.OracleModel(m, [dict(kind=orc.NEOHOOKEAN_DAMAGE, E=1e9, nu=0.3, rho=1000.0,
                                   Sd=1e4, Yd=np.full(m.nb_elements * m.nb_qp, 50.0))])
    cpu.nodal("v")[:] = v0
    cpu.blocked[:] = blocked
    cpu.set_time_step(model.time_step)
    cpu.solve_step(30)
    u_ref = cpu.nodal("u")
    assert np.abs(arr["displacement"] - u_ref).max() <= 1e-12 * np.abs(u_ref).max()
    assert np.abs(arr["velocity"] - cpu.nodal("v")).max() <= 1e-12 * np.abs(cpu.nodal("v")).max()
    dmg = cpu.field("damage").reshape(m.nb_elements, m.nb_qp).mean(axis=1)
    assert (dmg > 0).any()
    assert np.abs(arr["damage"] - dmg).max() <= 1e-12
    sig = cpu.field("stress").reshape(m.nb_elements, m.nb_qp, 9).mean(axis=1)
    assert np.abs(arr["stress"] - sig).max() <= 1e-12 * np.abs(sig).max()
    model.close()


def _hex_damage_model():
    from akantu_amd import akx
    m = akm.box_hex8(4, 3, 2)
    model = akx.SolidMechanicsModel(m)
    model.add_material(kind=akx.NEOHOOKEAN_DAMAGE, E=1e9, nu=0.3, rho=1000.0, Sd=1e4,
                       Yd=np.full(m.nb_elements * m.nb_qp, 50.0))
    model.init_full()
    rng = np.random.default_rng(1)
    model.set("velocity", rng.normal(scale=0.5, size=m.nodes.shape))
    blocked = np.zeros(m.nodes.shape, dtype=bool)
    blocked[m.nodes[:, 2] == 0.0] = True
    model.set("blocked_dofs", blocked)
    model.set_time_step(0.5 * model.get_stable_time_step())
    return model


ASYNC_FIELDS = ("displacement", "velocity", "acceleration", "internal_force", "mass",
                "blocked_dofs", "current_position", "grad_u", "stress", "damage")


@pytest.mark.gpu
def test_async_download_is_a_snapshot_of_the_enqueued_state():
    """Every asynchronous field equals the synchronous download taken at the
    same point, although more steps are enqueued before the copy is awaited."""
    from akantu_amd import akx
    model = _hex_damage_model()
    model.solve_step(30, akx.FP64)
    ref = {f: model.get(f) for f in ASYNC_FIELDS}
    bufs, tickets = {}, {}
    for f in ASYNC_FIELDS:
        shape, dtype = model.field_layout(f)
        bufs[f] = akx.PinnedBuffer(int(np.prod(shape)) * np.dtype(dtype).itemsize)
        tickets[f] = model.get_async(f, bufs[f])
    model.solve_step(30, akx.FP64)          # enqueued behind the snapshots
    for f in ASYNC_FIELDS:
        model.wait(tickets[f])
        assert model.done(tickets[f])
        shape, dtype = model.field_layout(f)
        got = bufs[f].array(dtype, shape)
        want = ref[f].astype(np.uint8) if f == "blocked_dofs" else ref[f]
        assert np.array_equal(got, want), f
    assert not np.array_equal(model.get_displacement(), ref["displacement"])
    with pytest.raises(akx.AkxError):
        model.get_async("jacobians", bufs["damage"])
    for b in bufs.values():
        b.close()
    model.close()


@pytest.mark.gpu
def test_async_dumps_equal_synchronous_dumps(tmp_path):
    """Dumps written by the writer thread while the model keeps stepping are
    the files the synchronous dumper writes at the same steps."""
    from akantu_amd import akx
    texts = {}
    for mode in (False, True):
        model = _hex_damage_model()
        model.set_base_name("h", str(tmp_path / str(mode)), asynchronous=mode)
        for f in ("displacement", "velocity", "blocked_dofs", "stress", "damage"):
            model.add_dump_field(f)
        model.dump()
        for _ in range(5):                  # more dumps than the pinned-buffer depth
            model.solve_step(10, akx.FP64)
            model.dump()
        model.flush_dumps()
        texts[mode] = [(tmp_path / str(mode) / ("h_%04d.vtu" % k)).read_text() for k in range(6)]
        texts[mode].append((tmp_path / str(mode) / "h.pvd").read_text())
        model.close()
    assert texts[True] == texts[False]
