"""ParaView output of a model: one VTK XML unstructured-grid file (.vtu) per
dump plus a .pvd collection that lists them with their time.

Plays the part of the reference's dumper for this path
(SolidMechanicsModel::setBaseName / addDumpField / dump,
src/model/model.cc and src/io/dumper/), which writes nodal fields
(displacement, velocity, ...) and per-element quadrature fields into
`paraview/<base>_<step>.vtu`.  Nodal fields are point data; per-point fields
of elements with several quadrature points are averaged per element as cell
data.  Arrays are written in VTK's inline binary format by default (per
DataArray one base64 stream of a UInt64 byte count followed by the raw
little-endian values, header_type="UInt64"), or as ASCII text with
binary=False.  The mesh part of a file (points, connectivity, offsets, cell
types) is encoded once per Dumper.

Asynchronous by default: dump() takes a device snapshot of every field
(akx_field_download_async), and one writer thread waits for the copies and
writes the files while the caller keeps stepping.  Page-locked buffers are
recycled once their file is written; at most `depth` dumps are in flight
(dump() waits for the oldest beyond that).  flush() waits for all of them.
"""
from __future__ import annotations

import base64
import collections
import concurrent.futures
import os
import threading
import xml.etree.ElementTree as ET

import numpy as np

from . import mesh as akm

# VTK cell types of the element types of the path
VTK_CELL = {akm.TRIANGLE_3: 5, akm.QUADRANGLE_4: 9, akm.TETRAHEDRON_4: 10, akm.HEXAHEDRON_8: 12}

NODAL = ("displacement", "velocity", "acceleration", "external_force", "internal_force",
         "mass", "blocked_dofs")
POINTWISE = ("stress", "grad_u", "damage")


_NP = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}


def _encode(values, kind, binary):
    """The text of one DataArray: base64(UInt64 byte count + raw data) or ASCII."""
    flat = np.ascontiguousarray(np.asarray(values).reshape(-1), dtype=_NP[kind])
    if binary:
        raw = flat.tobytes()
        return base64.b64encode(np.array([len(raw)], dtype="<u8").tobytes() + raw).decode("ascii")
    if kind == "Float64":
        return " ".join(repr(float(v)) for v in flat)
    return " ".join(str(int(v)) for v in flat)


def _add_array(parent, name, kind, ncomp, text, binary):
    da = ET.SubElement(parent, "DataArray", type=kind, Name=name,
                       NumberOfComponents=str(ncomp), format="binary" if binary else "ascii")
    da.text = text


def _array_xml(parent, name, values, ncomp, binary):
    values = np.asarray(values)
    kind = "UInt8" if values.dtype == np.uint8 or values.dtype == bool else "Float64"
    _add_array(parent, name, kind, ncomp, _encode(values, kind, binary), binary)


def mesh_arrays(mesh, binary=True):
    """The encoded mesh part of a .vtu (positions padded to 3D, connectivity,
    offsets, cell types), reusable for every dump of one mesh."""
    N, d = mesh.nodes.shape
    E, n = mesh.conn.shape
    xyz = np.zeros((N, 3))
    xyz[:, :d] = mesh.nodes
    return {"positions": _encode(xyz, "Float64", binary),
            "connectivity": _encode(mesh.conn, "Int64", binary),
            "offsets": _encode(n * np.arange(1, E + 1, dtype=np.int64), "Int64", binary),
            "types": _encode(np.full(E, VTK_CELL[mesh.elem_type], dtype=np.uint8), "UInt8",
                             binary),
            "binary": binary}


def write_vtu(path, mesh, point_data=None, cell_data=None, binary=True, mesh_text=None):
    """Write one .vtu: coordinates padded to 3D, point and cell arrays.
    mesh_text: the result of mesh_arrays(mesh, binary), to skip re-encoding."""
    N, d = mesh.nodes.shape
    E, n = mesh.conn.shape
    if mesh_text is None or mesh_text["binary"] != binary:
        mesh_text = mesh_arrays(mesh, binary)
    attrs = dict(type="UnstructuredGrid", version="0.1", byte_order="LittleEndian")
    if binary:
        attrs["header_type"] = "UInt64"
    root = ET.Element("VTKFile", **attrs)
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(N), NumberOfCells=str(E))
    pts = ET.SubElement(piece, "Points")
    _add_array(pts, "positions", "Float64", 3, mesh_text["positions"], binary)
    cells = ET.SubElement(piece, "Cells")
    _add_array(cells, "connectivity", "Int64", 1, mesh_text["connectivity"], binary)
    _add_array(cells, "offsets", "Int64", 1, mesh_text["offsets"], binary)
    _add_array(cells, "types", "UInt8", 1, mesh_text["types"], binary)
    if point_data:
        pd = ET.SubElement(piece, "PointData")
        for name, arr in point_data.items():
            arr = np.asarray(arr)
            ncomp = 1 if arr.ndim == 1 else arr.shape[1]
            if ncomp == 2:  # ParaView draws vectors with 3 components
                arr = np.concatenate([arr, np.zeros((arr.shape[0], 1), arr.dtype)], axis=1)
                ncomp = 3
            _array_xml(pd, name, arr, ncomp, binary)
    if cell_data:
        cd = ET.SubElement(piece, "CellData")
        for name, arr in cell_data.items():
            arr = np.asarray(arr)
            _array_xml(cd, name, arr, 1 if arr.ndim == 1 else arr.shape[1], binary)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def write_pvd(path, entries):
    """entries: (time, vtu file name relative to the .pvd)."""
    root = ET.Element("VTKFile", type="Collection", version="0.1")
    col = ET.SubElement(root, "Collection")
    for t, f in entries:
        ET.SubElement(col, "DataSet", timestep=repr(float(t)), group="", part="0", file=f)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


class Dumper:
    """model.set_base_name / add_dump_field / dump, for akx.SolidMechanicsModel."""

    def __init__(self, model, base_name, directory="paraview", asynchronous=True, depth=2,
                 binary=True):
        self.model = model
        self.binary = bool(binary)
        self._mesh_text = None
        self.base = base_name
        self.dir = directory
        self.fields = []
        self.count = 0
        self.entries = []
        self.asynchronous = asynchronous
        self.depth = max(1, int(depth))
        self._pool = None
        self._pending = collections.deque()
        self._free = collections.defaultdict(list)  # field -> idle PinnedBuffers
        self._lock = threading.Lock()

    def add_dump_field(self, name):
        if name not in NODAL + POINTWISE:
            raise ValueError("unknown dump field %r" % name)
        if name not in self.fields:
            self.fields.append(name)

    def _data(self, arrays):
        """point and cell data from the fields' arrays (as model.get returns them)."""
        m = self.model.mesh
        pd, cd = {}, {}
        for f, a in arrays:
            if f in NODAL:
                pd[f] = a.astype(np.uint8) if a.dtype == bool else a
            else:
                per = a.reshape(m.nb_elements, m.nb_qp, -1).mean(axis=1)
                cd[f] = per[:, 0] if per.shape[1] == 1 else per
        return pd, cd

    def _write(self, name, arrays, entries):
        if self._mesh_text is None:
            self._mesh_text = mesh_arrays(self.model.mesh, self.binary)
        write_vtu(os.path.join(self.dir, name), self.model.mesh, *self._data(arrays),
                  binary=self.binary, mesh_text=self._mesh_text)
        write_pvd(os.path.join(self.dir, self.base + ".pvd"), entries)

    def dump(self, time=None):
        name = "%s_%04d.vtu" % (self.base, self.count)
        t = self.count * self.model.time_step if time is None else time
        self.entries.append((t, name))
        self.count += 1
        path = os.path.join(self.dir, name)
        if not self.asynchronous:
            self._write(name, [(f, self.model.get(f)) for f in self.fields], list(self.entries))
            return path
        from .akx import PinnedBuffer
        while len(self._pending) >= self.depth:
            self._pending.popleft().result()
        jobs = []
        for f in self.fields:
            shape, dtype = self.model.field_layout(f)
            nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
            with self._lock:
                free = self._free[f]
                buf = free.pop() if free else None
            if buf is None or buf.nbytes < nbytes:
                buf = PinnedBuffer(nbytes)
            jobs.append((f, buf, self.model.get_async(f, buf), shape, dtype))
        if self._pool is None:
            self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        self._pending.append(self._pool.submit(self._write_async, name, jobs, list(self.entries)))
        return path

    def _write_async(self, name, jobs, entries):
        arrays = []
        for f, buf, ticket, shape, dtype in jobs:
            self.model.wait(ticket)
            a = buf.array(dtype, shape)
            arrays.append((f, a.astype(bool) if f == "blocked_dofs" else a))
        try:
            self._write(name, arrays, entries)
        finally:
            with self._lock:
                for f, buf, _, _, _ in jobs:
                    self._free[f].append(buf)

    def flush(self):
        """Wait until every dump so far is on disk (re-raises a writer error)."""
        while self._pending:
            self._pending.popleft().result()

    def close(self):
        try:
            self.flush()
        finally:
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
            with self._lock:
                for bufs in self._free.values():
                    for b in bufs:
                        b.close()
                self._free.clear()
