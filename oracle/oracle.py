"""ctypes binding of the CPU FP64 oracle (oracle/akx_oracle.cc).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or as the timed CPU
reference loop.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libakx_oracle.so")
_lib = None

ELASTIC, NEOHOOKEAN, NEOHOOKEAN_DAMAGE = 0, 1, 2
FIELDS = {"u": 0, "v": 1, "a": 2, "f_ext": 3, "f_int": 4, "mass": 5, "B": 6, "JxW": 7,
          "gradu": 8, "stress": 9, "damage": 10, "Yd": 11, "previous": 12,
          "increment": 13, "X": 14}
_D, _P, _I64 = ctypes.c_double, ctypes.c_void_p, ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def configure(L):
    """Set the ctypes signatures of an oracle library (also used for copies
    of the oracle built from deliberately altered sources, to show that the
    tests catch a planted mistake)."""
    L.orc_create.restype = _P
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _I64, _P, _I64, _P]
    L.orc_destroy.argtypes = [_P]
    L.orc_last_error.restype = ctypes.c_char_p
    L.orc_last_error.argtypes = [_P]
    L.orc_add_material.argtypes = [_P, ctypes.c_int, _D, _D, _D, ctypes.c_int]
    L.orc_set_damage.argtypes = [_P, ctypes.c_int, _D, _P]
    L.orc_assign_materials.argtypes = [_P, _P]
    L.orc_init.argtypes = [_P]
    L.orc_set_threads.argtypes = [_P, ctypes.c_int]
    L.orc_field.restype = ctypes.POINTER(_D)
    L.orc_field.argtypes = [_P, ctypes.c_int, ctypes.POINTER(_I64)]
    L.orc_blocked.restype = ctypes.POINTER(ctypes.c_uint8)
    L.orc_blocked.argtypes = [_P]
    L.orc_set_time_step.argtypes = [_P, _D]
    L.orc_assemble_internal_forces.argtypes = [_P]
    L.orc_solve_step.argtypes = [_P, ctypes.c_int]
    L.orc_stable_dt.restype = _D
    L.orc_stable_dt.argtypes = [_P]
    L.orc_energy.restype = _D
    L.orc_energy.argtypes = [_P, ctypes.c_int]
    L.orc_point_stress.restype = _D
    L.orc_point_stress.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int, _D, _D,
                                   _P, _P, _P]
    L.orc_point_energy.restype = _D
    L.orc_point_energy.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int, _P]
    L.orc_point_finite_residual.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
    L.orc_truss_dynamic_explicit.argtypes = [ctypes.c_char_p, _I64]
    for name in ("orc_fe_gradient", "orc_fe_interpolate", "orc_fe_btd", "orc_fe_ntb",
                 "orc_fe_integrate", "orc_fe_integrate_on_points"):
        getattr(L, name).argtypes = [_P, _P, ctypes.c_int, _P, _I64, _P]
    L.orc_fe_integrate_scalar.restype = _D
    L.orc_fe_integrate_scalar.argtypes = [_P, _P, _P, _I64]
    L.orc_fe_assemble_field_lumped.argtypes = [_P, _P, ctypes.c_int, _P]
    L.orc_material_compute_stress.argtypes = [_P, ctypes.c_int, _P, _P]
    L.orc_material_compute_all_stresses.argtypes = [_P, ctypes.c_int]
    L.orc_material_assemble_internal_forces.argtypes = [_P, ctypes.c_int]
    L.orc_material_potential_energy.restype = _D
    L.orc_material_potential_energy.argtypes = [_P, ctypes.c_int]
    L.orc_material_elements.restype = _I64
    L.orc_material_elements.argtypes = [_P, ctypes.c_int, _P]
    return L


def load(path):
    """An oracle library other than the default one, configured."""
    return configure(ctypes.CDLL(path))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = configure(ctypes.CDLL(_LIB_PATH))
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleModel:
    """FP64 CPU restatement of SolidMechanicsModel's explicit lumped-mass path."""

    def __init__(self, mesh, materials, elem_mat=None, threads=1, library=None):
        L = library if library is not None else lib()
        self._L = L
        self._X = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, dtype=np.uint32)
        self.dim = mesh.dim
        self.nb_nodes = mesh.nb_nodes
        self.nb_elements = mesh.nb_elements
        self.nb_nodes_per_element = mesh.conn.shape[1]
        self.nb_qp = mesh.nb_qp
        self.h = L.orc_create(mesh.elem_type, mesh.dim, mesh.nb_nodes, _ptr(self._X),
                              mesh.nb_elements, _ptr(self._conn))
        if not self.h:
            raise ValueError("oracle: bad mesh")
        for m in materials:
            L.orc_add_material(self.h, m.get("kind", ELASTIC), m["E"], m["nu"], m["rho"],
                               int(m.get("plane_stress", 0)))
        if elem_mat is not None:
            em = np.ascontiguousarray(elem_mat, dtype=np.int32)
            if L.orc_assign_materials(self.h, _ptr(em)) != 0:
                raise ValueError("oracle: bad material assignment")
        for k, m in enumerate(materials):
            if m.get("kind", ELASTIC) == NEOHOOKEAN_DAMAGE:
                yd = np.ascontiguousarray(m["Yd"], dtype=np.float64)
                L.orc_set_damage(self.h, k, m["Sd"], _ptr(yd))
        L.orc_set_threads(self.h, threads)
        rc = L.orc_init(self.h)
        if rc != 0:
            raise RuntimeError("oracle init failed: %s" % L.orc_last_error(self.h).decode())

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    def field(self, name):
        n = _I64()
        p = self._L.orc_field(self.h, FIELDS[name], ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def nodal(self, name):
        return self.field(name).reshape(self.nb_nodes, self.dim)

    @property
    def blocked(self):
        p = self._L.orc_blocked(self.h)
        return np.ctypeslib.as_array(p, shape=(self.nb_nodes * self.dim,)).reshape(
            self.nb_nodes, self.dim)

    def set_time_step(self, dt):
        self._L.orc_set_time_step(self.h, dt)

    def set_threads(self, n):
        self._L.orc_set_threads(self.h, n)

    def solve_step(self, nsteps=1):
        if self._L.orc_solve_step(self.h, nsteps) != 0:
            raise RuntimeError("oracle: not initialised")

    def assemble_internal_forces(self):
        self._L.orc_assemble_internal_forces(self.h)

    def stable_time_step(self):
        return self._L.orc_stable_dt(self.h)

    def energy(self, kind):
        return self._L.orc_energy(self.h, {"kinetic": 0, "potential": 1, "external work": 2}[kind])

    # -- FEEngine operators (fe_engine.hh:95-244), rows in `elements` order --
    def _count(self, elements):
        return self.nb_elements if elements is None else len(elements)

    def _fe(self, name, arr, nb, out_cols, per_point, elements):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        ids = None if elements is None else np.ascontiguousarray(elements, dtype=np.uint32)
        rows = self._count(elements) * (self.nb_qp if per_point else 1)
        out = np.zeros((rows, out_cols))
        rc = getattr(self._L, name)(self.h, _ptr(arr), int(nb),
                                    None if ids is None else _ptr(ids),
                                    0 if ids is None else ids.size, _ptr(out))
        if rc != 0:
            raise ValueError("oracle: %s rejected its arguments" % name)
        return out

    def fe_gradient(self, u, nb, elements=None):
        return self._fe("orc_fe_gradient", u, nb, nb * self.dim, True, elements)

    def fe_interpolate(self, u, nb, elements=None):
        return self._fe("orc_fe_interpolate", u, nb, nb, True, elements)

    def fe_btd(self, Ds, rows, elements=None):
        return self._fe("orc_fe_btd", Ds, rows, rows * self.nb_nodes_per_element, True, elements)

    def fe_ntb(self, bs, nb, elements=None):
        return self._fe("orc_fe_ntb", bs, nb, nb * self.nb_nodes_per_element, True, elements)

    def fe_integrate(self, f, nb, elements=None):
        return self._fe("orc_fe_integrate", f, nb, nb, False, elements)

    def fe_integrate_on_points(self, f, nb, elements=None):
        return self._fe("orc_fe_integrate_on_points", f, nb, nb, True, elements)

    def fe_integrate_scalar(self, f, elements=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        ids = None if elements is None else np.ascontiguousarray(elements, dtype=np.uint32)
        return self._L.orc_fe_integrate_scalar(self.h, _ptr(f), None if ids is None else _ptr(ids),
                                               0 if ids is None else ids.size)

    def fe_assemble_field_lumped(self, field, nb, lumped):
        field = np.ascontiguousarray(field, dtype=np.float64)
        out = np.ascontiguousarray(lumped, dtype=np.float64).copy()
        self._L.orc_fe_assemble_field_lumped(self.h, _ptr(field), int(nb), _ptr(out))
        return out

    # -- Material operators (material.hh:110-195) --
    def material_elements(self, mat):
        n = self._L.orc_material_elements(self.h, mat, None)
        ids = np.zeros(n, dtype=np.uint32)
        self._L.orc_material_elements(self.h, mat, _ptr(ids))
        return ids

    def material_compute_stress(self, mat, grad_u):
        g = np.ascontiguousarray(grad_u, dtype=np.float64)
        out = np.zeros_like(g)
        self._L.orc_material_compute_stress(self.h, mat, _ptr(g), _ptr(out))
        return out

    def material_compute_all_stresses(self, mat):
        self._L.orc_material_compute_all_stresses(self.h, mat)

    def material_assemble_internal_forces(self, mat):
        self._L.orc_material_assemble_internal_forces(self.h, mat)

    def material_potential_energy(self, mat):
        return self._L.orc_material_potential_energy(self.h, mat)


def point_stress(kind, dim, E, nu, G, plane_stress=0, Sd=0.0, Yd=0.0, damage=0.0):
    G = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(dim, dim).T)  # col-major
    S = np.zeros(dim * dim)
    d = np.array([damage], dtype=np.float64)
    lib().orc_point_stress(kind, dim, E, nu, plane_stress, Sd, Yd, _ptr(d), _ptr(G), _ptr(S))
    return S.reshape(dim, dim).T.copy(), float(d[0])


def point_energy(kind, dim, E, nu, G, plane_stress=0):
    G = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(dim, dim).T)
    return lib().orc_point_energy(kind, dim, E, nu, plane_stress, _ptr(G))


def point_finite_residual(dim, B, G, S):
    """B: (dim, n) matrix of dN_a/dx_i; G, S: (dim, dim).  Returns r (n, dim)."""
    n = B.shape[1]
    Bc = np.ascontiguousarray(np.asarray(B, dtype=np.float64).T.reshape(-1))  # col-major d x n
    Gc = np.ascontiguousarray(np.asarray(G, dtype=np.float64).T.reshape(-1))
    Sc = np.ascontiguousarray(np.asarray(S, dtype=np.float64).T.reshape(-1))
    r = np.zeros(dim * n)
    lib().orc_point_finite_residual(dim, n, _ptr(Bc), _ptr(Gc), _ptr(Sc), _ptr(r))
    return r.reshape(n, dim)


def truss_dynamic_explicit() -> str:
    buf = ctypes.create_string_buffer(1 << 16)
    n = lib().orc_truss_dynamic_explicit(buf, len(buf))
    if n < 0:
        raise RuntimeError("buffer too small")
    return buf.value.decode()
