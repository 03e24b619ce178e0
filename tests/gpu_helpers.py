"""Set-up shared by the GPU parity tests: the same mesh, materials, boundary
conditions and initial state handed to the GPU model (through the C ABI) and
to the CPU FP64 oracle."""
import numpy as np

from akantu_amd import akx
from akantu_amd import mesh as akm
from oracle import oracle as orc


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


class Pair:
    """A GPU model and the oracle model, kept in lock-step."""

    def __init__(self, mesh, materials, elem_mat=None, keep_numbering=False):
        self.mesh = mesh
        self.gpu = akx.SolidMechanicsModel(mesh, keep_numbering=keep_numbering)
        for m in materials:
            self.gpu.add_material(kind=m.get("kind", akx.ELASTIC), E=m["E"], nu=m["nu"],
                                  rho=m["rho"], plane_stress=m.get("plane_stress", 0),
                                  Sd=m.get("Sd"), Yd=m.get("Yd"))
        if elem_mat is not None:
            self.gpu.assign_materials(elem_mat)
        self.gpu.init_full()
        self.cpu = orc.OracleModel(mesh, materials, elem_mat)

    def set_nodal(self, name, values):
        oname = {"displacement": "u", "velocity": "v", "acceleration": "a",
                 "external_force": "f_ext"}[name]
        self.gpu.set(name, values)
        self.cpu.nodal(oname)[:] = values

    def block(self, mask):
        mask = np.asarray(mask, dtype=bool)
        self.gpu.set("blocked_dofs", mask)
        self.cpu.blocked[:] = mask

    def set_time_step(self, dt):
        self.gpu.set_time_step(dt)
        self.cpu.set_time_step(dt)

    def step(self, n, precision=akx.FP64):
        self.gpu.solve_step(n, precision)
        self.cpu.solve_step(n)


def bar_pair(nx=100, ny=20):
    """c0: 1.0 x 0.2 m plane-strain bar of triangle_3, left edge fixed,
    v_x = 0.01 sin(pi x / 2), dt = 0.8 x stable."""
    m = akm.bar_tri3(nx, ny)
    p = Pair(m, [dict(E=210e9, nu=0.3, rho=7800.0)])
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 0], 0.0)] = True
    p.block(mask)
    v = np.zeros_like(X)
    v[:, 0] = 0.01 * np.sin(np.pi * X[:, 0] / 2)
    p.set_nodal("velocity", v)
    dt = 0.8 * p.cpu.stable_time_step()
    p.set_time_step(dt)
    return p


def cube_pair(n=10, jitter=0.01, seed=1):
    """reduced c1: unit cube of n^3 Kuhn-split hexes (tetrahedron_4), interior
    jitter, bottom face blocked, top face v_z = -1 m/s."""
    m = akm.cube_tet4(n, jitter, seed)
    p = Pair(m, [dict(E=210e9, nu=0.3, rho=7800.0)])
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 2], 0.0)] = True
    p.block(mask)
    v = np.zeros_like(X)
    v[np.isclose(X[:, 2], 1.0), 2] = -1.0
    p.set_nodal("velocity", v)
    p.set_time_step(0.8 * p.cpu.stable_time_step())
    return p


def hex_pair(n=6, kind=akx.NEOHOOKEAN, Yd=None, Sd=None, speed=0.5):
    """reduced c2: unit cube of n^3 hexahedron_8, neo-Hookean E=10 MPa nu=.45
    rho=1000, bottom blocked, top face pushed down at `speed` m/s through a
    per-step Dirichlet increment."""
    m = akm.cube_hex8(n)
    mat = dict(kind=kind, E=1e7, nu=0.45, rho=1000.0)
    if kind == akx.NEOHOOKEAN_DAMAGE:
        mat.update(Sd=Sd, Yd=Yd)
    p = Pair(m, [mat])
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 2], 0.0)] = True
    p.block(mask)
    dt = 0.8 * p.cpu.stable_time_step()
    p.set_time_step(dt)
    top = np.where(np.isclose(X[:, 2], 1.0))[0]
    inc = -speed * dt
    p.gpu.set_dirichlet_increment(top, 2, inc)
    p.top, p.inc = top, inc
    return p


def oracle_hex_step(p, nsteps):
    """Advance the oracle of hex_pair with the same per-step increment
    (applyBC(IncrementValue) before every solveStep)."""
    u = p.cpu.nodal("u")
    p.cpu.blocked[p.top, 2] = 1
    for _ in range(nsteps):
        u[p.top, 2] += p.inc
        p.cpu.solve_step(1)
