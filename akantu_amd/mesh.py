"""Host-side meshes: gmsh 2.x reader, structured generators for the explicit
workloads, and locality renumbering.

Conventions follow the reference ``Mesh`` (src/mesh/mesh.hh:100-611):
nodes are an (N, d) row-major float64 array, connectivity an (E, n) uint32
array with the reference's local node order for each element type
(element_class_*_inline_impl.hh).  Element type ids are the C-ABI ids of
``include/akx.h``.
"""
from __future__ import annotations

import dataclasses
import itertools
import math

import numpy as np

TRIANGLE_3, QUADRANGLE_4, TETRAHEDRON_4, HEXAHEDRON_8 = 0, 1, 2, 3

ELEMENT_DIM = {TRIANGLE_3: 2, QUADRANGLE_4: 2, TETRAHEDRON_4: 3, HEXAHEDRON_8: 3}
ELEMENT_NODES = {TRIANGLE_3: 3, QUADRANGLE_4: 4, TETRAHEDRON_4: 4, HEXAHEDRON_8: 8}
ELEMENT_QP = {TRIANGLE_3: 1, QUADRANGLE_4: 4, TETRAHEDRON_4: 1, HEXAHEDRON_8: 8}
ELEMENT_NAMES = {TRIANGLE_3: "triangle_3", QUADRANGLE_4: "quadrangle_4",
                 TETRAHEDRON_4: "tetrahedron_4", HEXAHEDRON_8: "hexahedron_8"}

# gmsh 2.x element codes -> element type (src/io/mesh_io/mesh_io_msh.cc:181-199);
# the four linear types use gmsh's node order unchanged (:229-275).
_MSH_TYPES = {2: TRIANGLE_3, 3: QUADRANGLE_4, 4: TETRAHEDRON_4, 5: HEXAHEDRON_8}


@dataclasses.dataclass
class Mesh:
    """One element type, one spatial dimension."""

    elem_type: int
    nodes: np.ndarray  # (N, d) float64
    conn: np.ndarray  # (E, n) uint32
    groups: dict = dataclasses.field(default_factory=dict)  # name -> node ids

    @property
    def dim(self) -> int:
        return ELEMENT_DIM[self.elem_type]

    @property
    def nb_nodes(self) -> int:
        return self.nodes.shape[0]

    @property
    def nb_elements(self) -> int:
        return self.conn.shape[0]

    @property
    def nb_qp(self) -> int:
        return ELEMENT_QP[self.elem_type]

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.float64)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.uint32)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != ELEMENT_DIM[self.elem_type]:
            raise ValueError("nodes must be (N, %d)" % ELEMENT_DIM[self.elem_type])
        if self.conn.ndim != 2 or self.conn.shape[1] != ELEMENT_NODES[self.elem_type]:
            raise ValueError("conn must be (E, %d)" % ELEMENT_NODES[self.elem_type])
        if self.conn.size and int(self.conn.max()) >= self.nodes.shape[0]:
            raise ValueError("connectivity refers to a missing node")


def read_msh(path: str, dim: int | None = None) -> Mesh:
    """Read a gmsh 2.x ASCII file (MeshIOMSH::read, mesh_io_msh.cc).

    Keeps the elements of the highest dimension present (or of ``dim``);
    node ids are renumbered in order of appearance, coordinates are truncated
    to the spatial dimension.  Physical tags become node groups "phys_<tag>".
    """
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    i = 0
    ids, coords, elems = [], [], []
    while i < len(lines):
        tag = lines[i]
        if tag == "$Nodes":
            n = int(lines[i + 1])
            for k in range(n):
                p = lines[i + 2 + k].split()
                ids.append(int(p[0]))
                coords.append([float(v) for v in p[1:4]])
            i += n + 3
            continue
        if tag == "$Elements":
            n = int(lines[i + 1])
            for k in range(n):
                p = [int(v) for v in lines[i + 2 + k].split()]
                etype, ntags = p[1], p[2]
                tags = p[3:3 + ntags]
                nodes = p[3 + ntags:]
                elems.append((etype, tags[0] if tags else 0, nodes))
            i += n + 3
            continue
        i += 1
    id_map = {gid: k for k, gid in enumerate(ids)}
    candidates = [(ELEMENT_DIM[_MSH_TYPES[t]], _MSH_TYPES[t])
                  for t, _, _ in elems if t in _MSH_TYPES]
    if not candidates:
        raise ValueError("no supported elements in %s" % path)
    if dim is None:
        dim = max(c[0] for c in candidates)
    types = sorted({c[1] for c in candidates if c[0] == dim})
    if len(types) != 1:
        raise ValueError("expected exactly one element type of dimension %d" % dim)
    etype = types[0]
    code = [k for k, v in _MSH_TYPES.items() if v == etype][0]
    conn = np.array([[id_map[g] for g in nodes] for t, _, nodes in elems if t == code],
                    dtype=np.uint32)
    groups = {}
    for t, phys, nodes in elems:
        if phys:
            groups.setdefault("phys_%d" % phys, set()).update(id_map[g] for g in nodes)
    groups = {k: np.array(sorted(v), dtype=np.int64) for k, v in groups.items()}
    X = np.array(coords, dtype=np.float64)[:, :dim]
    return Mesh(etype, X, conn, groups)


# ----------------------------------------------------------------------------
# facets / boundary
# ----------------------------------------------------------------------------
_FACETS = {
    TRIANGLE_3: [(0, 1), (1, 2), (2, 0)],
    QUADRANGLE_4: [(0, 1), (1, 2), (2, 3), (3, 0)],
    TETRAHEDRON_4: [(0, 2, 1), (1, 2, 3), (2, 0, 3), (0, 1, 3)],
    HEXAHEDRON_8: [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5),
                   (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)],
}


def write_msh(mesh: Mesh, path: str, groups: dict | None = None) -> None:
    """Write a gmsh 2.2 ASCII file that read_msh (and Mesh::read of the C++
    API) reads back: 1-based node ids, coordinates padded to 3D, one
    element block of the mesh's type with physical tag 0, and, per entry
    ``tag -> node ids`` of ``groups``, one gmsh point element (type 15) per node
    with that physical tag, which the readers turn into node group
    "phys_<tag>".
    """
    code = [k for k, v in _MSH_TYPES.items() if v == mesh.elem_type][0]
    N, d = mesh.nodes.shape
    xyz = np.zeros((N, 3))
    xyz[:, :d] = mesh.nodes
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(N)]
    lines += ["%d %.17g %.17g %.17g" % (k + 1, x, y, z) for k, (x, y, z) in enumerate(xyz)]
    lines += ["$EndNodes", "$Elements"]
    body = []
    for e, row in enumerate(mesh.conn):
        body.append("%d %d 2 0 0 %s" % (len(body) + 1, code, " ".join(str(int(v) + 1) for v in row)))
    for tag, nodes in (groups or {}).items():
        for v in nodes:
            body.append("%d 15 2 %d %d %d" % (len(body) + 1, int(tag), int(tag), int(v) + 1))
    lines += [str(len(body))] + body + ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def boundary_facets(mesh: Mesh, with_owner: bool = False):
    """Facets (node ids) that belong to one element only; optionally also the
    owning element of each facet."""
    faces, owner = [], []
    E = mesh.nb_elements
    for loc in _FACETS[mesh.elem_type]:
        faces.append(mesh.conn[:, list(loc)])
        owner.append(np.arange(E))
    allf = np.concatenate(faces, axis=0).astype(np.int64)
    own = np.concatenate(owner)
    key = np.sort(allf, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    sel = cnt[inv] == 1
    if with_owner:
        return allf[sel], own[sel]
    return allf[sel]


def boundary_nodes(mesh: Mesh) -> np.ndarray:
    return np.unique(boundary_facets(mesh).reshape(-1))


# ----------------------------------------------------------------------------
# structured generators
# ----------------------------------------------------------------------------
def _grid_nodes(shape, lengths, origin=None):
    axes = [np.linspace(0.0, L, n + 1) for n, L in zip(shape, lengths)]
    if origin is not None:
        axes = [a + o for a, o in zip(axes, origin)]
    mesh = np.meshgrid(*axes, indexing="ij")
    # x fastest: node id = i + (nx+1) * (j + (ny+1) * k)
    return np.stack([m.transpose(*reversed(range(len(shape)))).reshape(-1) for m in mesh], axis=1)


def rectangle_tri3(nx: int, ny: int, lx: float, ly: float, origin=(0.0, 0.0)) -> Mesh:
    """lx x ly rectangle, nx x ny quads each split into 2 triangle_3 along the
    (i,j)-(i+1,j+1) diagonal; nodes and quads numbered x fastest."""
    X = _grid_nodes((nx, ny), (lx, ly), origin)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i, j = i.reshape(-1), j.reshape(-1)
    n00 = j * (nx + 1) + i
    n10, n01, n11 = n00 + 1, n00 + nx + 1, n00 + nx + 2
    conn = np.empty((2 * nx * ny, 3), dtype=np.uint32)
    conn[0::2] = np.stack([n00, n10, n11], axis=1)
    conn[1::2] = np.stack([n00, n11, n01], axis=1)
    return Mesh(TRIANGLE_3, X, conn)


def rectangle_quad4(nx, ny, lx, ly) -> Mesh:
    X = _grid_nodes((nx, ny), (lx, ly))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    n00 = (j * (nx + 1) + i).reshape(-1)
    conn = np.stack([n00, n00 + 1, n00 + nx + 2, n00 + nx + 1], axis=1)
    return Mesh(QUADRANGLE_4, X, conn)


def _hex_corners(nx, ny, nz, keep=None):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).reshape(-1) for a in (i, j, k))  # x fastest
    if keep is not None:
        m = keep(i, j, k)
        i, j, k = i[m], j[m], k[m]
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    base = i * sx + j * sy + k * sz
    # reference hexahedron_8 order (-,-,-), (+,-,-), (+,+,-), (-,+,-), then z+
    off = [0, sx, sx + sy, sy, sz, sz + sx, sz + sx + sy, sz + sy]
    return np.stack([base + o for o in off], axis=1)


def box_hex8(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0) -> Mesh:
    X = _grid_nodes((nx, ny, nz), (lx, ly, lz))
    return Mesh(HEXAHEDRON_8, X, _hex_corners(nx, ny, nz))


# Kuhn split of a unit cube into 6 tetrahedra along the (0,0,0)-(1,1,1)
# diagonal: one tet per axis permutation, vertices along the monotone path.
def _kuhn_tets():
    corner = {(0, 0, 0): 0, (1, 0, 0): 1, (1, 1, 0): 2, (0, 1, 0): 3,
              (0, 0, 1): 4, (1, 0, 1): 5, (1, 1, 1): 6, (0, 1, 1): 7}
    tets = []
    for perm in itertools.permutations(range(3)):
        p = [0, 0, 0]
        verts = [corner[tuple(p)]]
        for ax in perm:
            p[ax] = 1
            verts.append(corner[tuple(p)])
        inv = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b])
        if inv % 2 == 1:  # odd permutation: negative orientation, swap two vertices
            verts[2], verts[3] = verts[3], verts[2]
        tets.append(verts)
    return np.array(tets, dtype=np.int64)


def box_tet4(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, keep=None, origin=None) -> Mesh:
    """Structured box of nx*ny*nz hexes, each Kuhn-split into 6 positively
    oriented tetrahedron_4 (6 consecutive elements per hex, hexes x fastest)."""
    X = _grid_nodes((nx, ny, nz), (lx, ly, lz), origin)
    hexes = _hex_corners(nx, ny, nz, keep)
    kt = _kuhn_tets()
    conn = hexes[:, kt].reshape(-1, 4)
    return Mesh(TETRAHEDRON_4, X, conn)


def jitter_interior(mesh: Mesh, amplitude: float, seed: int) -> Mesh:
    """Uniform +-amplitude perturbation of every coordinate of the nodes that are
    not on the bounding box, with numpy's PCG64 seeded by ``seed``."""
    X = mesh.nodes.copy()
    lo, hi = X.min(axis=0), X.max(axis=0)
    tol = 1e-12 * max(1.0, float(np.abs(hi - lo).max()))
    interior = np.all((X > lo + tol) & (X < hi - tol), axis=1)
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-amplitude, amplitude, size=X.shape)
    X[interior] += noise[interior]
    return Mesh(mesh.elem_type, X, mesh.conn, dict(mesh.groups))


def unused_nodes_removed(mesh: Mesh) -> Mesh:
    used = np.zeros(mesh.nb_nodes, dtype=bool)
    used[mesh.conn.reshape(-1)] = True
    new_id = np.cumsum(used) - 1
    return Mesh(mesh.elem_type, mesh.nodes[used], new_id[mesh.conn].astype(np.uint32))


# ----------------------------------------------------------------------------
# the five explicit workloads
# ----------------------------------------------------------------------------
def bar_tri3(nx=100, ny=20, lx=1.0, ly=0.2) -> Mesh:
    """2D plane-strain bar 1.0 x 0.2 m, 100 x 20 quads -> 4,000 triangle_3, 2,121 nodes."""
    return rectangle_tri3(nx, ny, lx, ly)


def cube_tet4(n=100, jitter=0.01, seed=1) -> Mesh:
    """Unit cube, n^3 hexes Kuhn-split into 6 n^3 tetrahedron_4, interior nodes
    jittered by +-jitter*h (h = 1/n)."""
    m = box_tet4(n, n, n)
    if jitter:
        m = jitter_interior(m, jitter / n, seed)
    return m


def cube_hex8(n=128) -> Mesh:
    return box_hex8(n, n, n)


def notched_plate_tet4(nx=150, ny=150, nz=30, lx=0.1, ly=0.1, lz=0.02,
                       notch_len=0.3, notch_width=2) -> Mesh:
    """Plate lx x ly x lz with an edge notch: a slot of ``notch_width`` cells at
    mid-height entering from x = 0 over ``notch_len`` of the width."""
    jmid = ny // 2
    ilen = int(round(notch_len * nx))

    def keep(i, j, k):
        in_slot = (i < ilen) & (j >= jmid - notch_width // 2) & (j < jmid + (notch_width + 1) // 2)
        return ~in_slot

    return unused_nodes_removed(box_tet4(nx, ny, nz, lx, ly, lz, keep=keep))


@dataclasses.dataclass
class Specimen:
    mesh: Mesh
    E: float
    nu: float
    rho: float
    velocity: np.ndarray  # (N, 2) initial velocity
    blocked: np.ndarray  # (N, 2) bool


def ensemble_tri3(count=512, min_elems=2000, max_elems=40000, seed=2024):
    """Independent plane-strain triangle_3 specimens on seeded rectangular meshes.

    Element count uniform in [min_elems, max_elems]; aspect ratio in [1, 4];
    E ~ U[50, 300] GPa, nu ~ U[0.2, 0.35], rho ~ U[2000, 8000] kg/m^3; left
    edge blocked; initial velocity v = A (sin(pi m x/L) , sin(pi k y/H)) scaled
    to at most 0.05 m/s."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        target = rng.uniform(min_elems, max_elems)
        aspect = rng.uniform(1.0, 4.0)
        ny = max(2, int(round(math.sqrt(target / (2.0 * aspect)))))
        nx = max(2, int(round(target / (2.0 * ny))))
        L = rng.uniform(0.5, 2.0)
        H = L * ny / nx
        m = rectangle_tri3(nx, ny, L, H)
        E = rng.uniform(50e9, 300e9)
        nu = rng.uniform(0.2, 0.35)
        rho = rng.uniform(2000.0, 8000.0)
        amp = rng.uniform(0.005, 0.05, size=2)
        mx, ky = rng.integers(1, 5, size=2)
        X = m.nodes
        v = np.stack([amp[0] * np.sin(np.pi * mx * X[:, 0] / (2 * L)),
                      amp[1] * np.sin(np.pi * ky * X[:, 1] / H)], axis=1)
        blocked = np.zeros_like(X, dtype=bool)
        left = np.isclose(X[:, 0], 0.0)
        blocked[left] = True
        v[left] = 0.0
        out.append(Specimen(m, E, nu, rho, v, blocked))
    return out


@dataclasses.dataclass
class Ensemble:
    """Independent specimens meshed into one element-disjoint mesh."""
    mesh: Mesh
    node_offsets: np.ndarray  # (S+1,) first node of each specimen
    elem_offsets: np.ndarray  # (S+1,) first element of each specimen
    element_material: np.ndarray  # (E,) specimen index of every element
    velocity: np.ndarray  # (N, d)
    blocked: np.ndarray  # (N, d) bool
    specimens: list

    def node_specimen(self) -> np.ndarray:
        return np.repeat(np.arange(len(self.specimens)), np.diff(self.node_offsets))

    def specimen_mesh(self, k: int) -> Mesh:
        """Specimen k as a mesh of its own, with the coordinates it has in the
        ensemble (so that a model of it alone sees the same numbers)."""
        n0, n1 = self.node_offsets[k], self.node_offsets[k + 1]
        e0, e1 = self.elem_offsets[k], self.elem_offsets[k + 1]
        return Mesh(self.mesh.elem_type, self.mesh.nodes[n0:n1],
                    self.mesh.conn[e0:e1].astype(np.int64) - n0)


def concatenate_specimens(specimens, gap: float = 0.1) -> Ensemble:
    """Lay the specimens out on a square grid of tiles (no two overlap) and
    concatenate nodes and elements in specimen order: specimen k owns nodes
    node_offsets[k]:node_offsets[k+1] and elements elem_offsets[k]:elem_offsets[k+1]."""
    S = len(specimens)
    cols = max(1, int(math.ceil(math.sqrt(S))))
    ext = np.array([sp.mesh.nodes.max(axis=0) - sp.mesh.nodes.min(axis=0) for sp in specimens])
    tile = ext.max(axis=0) + gap
    nodes, conn, vel, blk, em = [], [], [], [], []
    noff, eoff = [0], [0]
    for k, sp in enumerate(specimens):
        shift = np.zeros(sp.mesh.dim)
        shift[0] = (k % cols) * tile[0]
        shift[1] = (k // cols) * tile[1]
        nodes.append(sp.mesh.nodes + shift)
        conn.append(sp.mesh.conn.astype(np.int64) + noff[-1])
        vel.append(sp.velocity)
        blk.append(sp.blocked)
        em.append(np.full(sp.mesh.nb_elements, k, dtype=np.int32))
        noff.append(noff[-1] + sp.mesh.nb_nodes)
        eoff.append(eoff[-1] + sp.mesh.nb_elements)
    m = Mesh(specimens[0].mesh.elem_type, np.concatenate(nodes), np.concatenate(conn))
    return Ensemble(m, np.array(noff), np.array(eoff), np.concatenate(em),
                    np.concatenate(vel), np.concatenate(blk), list(specimens))


# ----------------------------------------------------------------------------
# locality renumbering
# ----------------------------------------------------------------------------
def _morton3(ix, iy, iz, bits=21):
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v
    return spread(ix) | (spread(iy) << np.uint64(1)) | (spread(iz) << np.uint64(2))


def renumber_morton(mesh: Mesh):
    """Renumber nodes and elements along a Morton (Z-order) curve of their
    positions / centroids, so that elements sharing a node sit close together
    in memory.  Returns (new_mesh, node_perm, elem_perm) where
    ``new.nodes = old.nodes[node_perm]`` and ``new.conn = inv[old.conn[elem_perm]]``."""
    X = mesh.nodes
    lo, hi = X.min(axis=0), X.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)

    def key(P):
        q = np.clip(((P - lo) / span * (2 ** 20 - 1)).astype(np.int64), 0, 2 ** 20 - 1)
        if P.shape[1] == 2:
            q = np.concatenate([q, np.zeros((q.shape[0], 1), dtype=np.int64)], axis=1)
        return _morton3(q[:, 0], q[:, 1], q[:, 2])

    node_perm = np.argsort(key(X), kind="stable")
    inv = np.empty_like(node_perm)
    inv[node_perm] = np.arange(node_perm.size)
    cent = X[mesh.conn.astype(np.int64)].mean(axis=1)
    elem_perm = np.argsort(key(cent), kind="stable")
    conn = inv[mesh.conn[elem_perm].astype(np.int64)].astype(np.uint32)
    new = Mesh(mesh.elem_type, X[node_perm], conn,
               {k: inv[v] for k, v in mesh.groups.items()})
    return new, node_perm, elem_perm
