"""ctypes binding of libakx.so (the C ABI of include/akx.h).

`SolidMechanicsModel` mirrors the reference class of the same name
(src/model/solid_mechanics/solid_mechanics_model.hh:58-543) for the explicit
lumped-mass path: the constructor takes a mesh, materials are added as in the
reference's material file, `init_full()` corresponds to
``initFull(_analysis_method = _explicit_lumped_mass)``, and the nodal arrays
are reached through get_/set_ accessors that copy to and from the GPU.

There is no CPU fallback: importing this module raises if libakx.so was not
built, and every call raises AkxError on a non-zero status.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libakx.so")

TRIANGLE_3, QUADRANGLE_4, TETRAHEDRON_4, HEXAHEDRON_8 = 0, 1, 2, 3
ELASTIC, NEOHOOKEAN, NEOHOOKEAN_DAMAGE = 0, 1, 2
FP64, MIXED = 0, 1
SURROGATE_FORMATS = {"fp16": 0, "bf16": 1}
FIELDS = {"displacement": 0, "velocity": 1, "acceleration": 2, "external_force": 3,
          "internal_force": 4, "mass": 5, "blocked_dofs": 6, "previous_displacement": 7,
          "displacement_increment": 8, "grad_u": 9, "stress": 10, "damage": 11,
          "shapes_derivatives": 12, "jacobians": 13, "current_position": 14}
ENERGIES = {"kinetic": 0, "potential": 1, "external work": 2}

# Every entry point of include/akx.h, with its ctypes signature.
_P, _I, _I64, _D, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
_PD = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "akx_ctx_create": (_I, [_I, ctypes.c_uint, ctypes.POINTER(_P)]),
    "akx_ctx_destroy": (_I, [_P]),
    "akx_last_error": (ctypes.c_char_p, [_P]),
    "akx_set_stream": (_I, [_P, _P]),
    "akx_mesh_set": (_I, [_P, _I, _I, _I64, _P, _I64, _P]),
    "akx_material_add": (_I, [_P, _I, _D, _D, _D, _I, ctypes.POINTER(_I)]),
    "akx_material_damage": (_I, [_P, _I, _D, _P]),
    "akx_assign_materials": (_I, [_P, _P]),
    "akx_init": (_I, [_P]),
    "akx_field_upload": (_I, [_P, _I, _P, _SZ]),
    "akx_field_download": (_I, [_P, _I, _P, _SZ]),
    "akx_host_alloc": (_I, [_SZ, ctypes.POINTER(_P)]),
    "akx_host_free": (_I, [_P]),
    "akx_field_download_async": (_I, [_P, _I, _P, _SZ, ctypes.POINTER(ctypes.c_uint64)]),
    "akx_download_query": (_I, [_P, ctypes.c_uint64, ctypes.POINTER(_I)]),
    "akx_download_wait": (_I, [_P, ctypes.c_uint64]),
    "akx_field_device_pointer": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "akx_node_numbering": (_I, [_P, _P]),
    "akx_set_time_step": (_I, [_P, _D]),
    "akx_solve_step": (_I, [_P, _I, _I]),
    "akx_solve_step_graph": (_I, [_P, _I, _I]),
    "akx_set_node_time_steps": (_I, [_P, _P]),
    "akx_stable_time_step_per_material": (_I, [_P, _P]),
    "akx_assemble_internal_forces": (_I, [_P, _I]),
    "akx_stable_time_step": (_I, [_P, _PD]),
    "akx_energy": (_I, [_P, _I, _PD]),
    "akx_set_dirichlet_increment": (_I, [_P, _I64, _P, _I, _D]),
    "akx_synchronize": (_I, [_P]),
    "akx_enable_timing": (_I, [_P, _I]),
    "akx_last_step_timing": (_I, [_P, _PD, _PD]),
    "akx_set_launch": (_I, [_P, _I, _I, _I, _I, _I, _I]),
    "akx_autotune": (_I, [_P, _I, _I, _P, _PD]),
    "akx_query": (_I, [_P, ctypes.c_char_p, _PD]),
    "akx_surrogate_setup": (_I, [_P, _I, _D]),
    "akx_surrogate_generate": (_I, [_P, _I64, _I, _D, _D, _D, ctypes.c_uint64]),
    "akx_surrogate_train": (_I, [_P, _I, _I, _D, ctypes.c_uint64, _PD]),
    "akx_surrogate_params": (_I, [_P, _P, _I64, _I]),
    "akx_surrogate_gradients": (_I, [_P, _I64, _P, _P, _P, _PD]),
    "akx_surrogate_infer": (_I, [_P, _I64, _P, _P]),
    "akx_surrogate_dataset": (_I, [_P, _I64, _I64, _P, _P]),
    "akx_surrogate_enable": (_I, [_P, _I]),
    "akx_surrogate_fallbacks": (_I, [_P, ctypes.POINTER(ctypes.c_uint64), _I]),
    "akx_surrogate_format": (_I, [_P, _I]),
    "akx_fe_gradient_on_integration_points": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_interpolate_on_integration_points": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_compute_btd": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_compute_ntb": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_integrate": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_integrate_scalar": (_I, [_P, _P, _P, _I64, _PD]),
    "akx_fe_integrate_on_integration_points": (_I, [_P, _P, _I, _P, _I64, _P]),
    "akx_fe_assemble_field_lumped": (_I, [_P, _P, _I, _P]),
    "akx_material_elements": (_I, [_P, _I, ctypes.POINTER(_I64), _P]),
    "akx_material_compute_stress": (_I, [_P, _I, _P, _P]),
    "akx_material_compute_all_stresses": (_I, [_P, _I, _P, _P]),
    "akx_material_assemble_internal_forces": (_I, [_P, _I]),
    "akx_material_potential_energy": (_I, [_P, _I, _PD]),
    "akx_energy_noalloc": (_I, [_I, _I, _I64, _P, _P, _P, _P, _P, _D, _P, _P]),
}


class AkxError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libakx.so is not built (run __graft_entry__.build() or "
                          "python akantu_amd/build.py); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class PinnedBuffer:
    """Page-locked host memory (akx_host_alloc) for asynchronous downloads."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        if lib.akx_host_alloc(self.nbytes, ctypes.byref(p)) != 0:
            raise AkxError("akx_host_alloc(%d) failed" % self.nbytes)
        self.ptr = p

    def array(self, dtype, shape):
        """numpy view of the buffer (valid while the buffer lives)."""
        dtype = np.dtype(dtype)
        count = int(np.prod(shape))
        if count * dtype.itemsize > self.nbytes:
            raise ValueError("view larger than the buffer")
        raw = (ctypes.c_char * (count * dtype.itemsize)).from_address(self.ptr.value)
        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape)

    def close(self):
        if self.ptr:
            lib.akx_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class Context:
    """Owns one akx_ctx (device buffers + stream)."""

    def __init__(self, device: int = 0, keep_numbering: bool = False):
        h = _P()
        rc = lib.akx_ctx_create(device, 1 if keep_numbering else 0, ctypes.byref(h))
        if rc != 0:
            raise AkxError("akx_ctx_create failed (%d): no usable HIP device %d" % (rc, device))
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise AkxError("akx error %d: %s" % (rc, lib.akx_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            lib.akx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _elements(elements):
    if elements is None:
        return None, 0, None
    ids = np.ascontiguousarray(elements, dtype=np.uint32)
    return _ptr(ids), ids.size, ids


class FEEngine:
    """FEEngine operators of a model's mesh (src/fe_engine/fe_engine.hh:95-244),
    FP64 on the GPU in the reference's operation order.  `elements` is the
    reference's filter_elements (caller element ids; None = all); per-point
    arrays have one row per point of the listed elements (element-major)."""

    def __init__(self, model):
        self.model = model
        self.ctx = model.ctx
        m = model.mesh
        self.dim, self.nb_nodes_per_element, self.nb_qp = m.dim, m.conn.shape[1], m.nb_qp

    def _count(self, elements):
        return self.model.mesh.nb_elements if elements is None else len(elements)

    def get_nb_integration_points(self):
        return self.nb_qp

    def get_shapes_derivatives(self):
        return self.model.get("shapes_derivatives")

    def get_jacobians(self):
        return self.model.get("jacobians")

    def _op(self, fn, arr, nb, out_cols, elements):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p, n, keep = _elements(elements)
        out = np.zeros((self._count(elements) * self.nb_qp, out_cols))
        self.ctx.check(fn(self.ctx.h, _ptr(arr), int(nb), p, n, _ptr(out)))
        return out

    def gradient_on_integration_points(self, u, nb_degree_of_freedom=None, elements=None):
        u = np.asarray(u, dtype=np.float64)
        nb = nb_degree_of_freedom or (u.shape[1] if u.ndim == 2 else 1)
        return self._op(lib.akx_fe_gradient_on_integration_points, u, nb, nb * self.dim, elements)

    def interpolate_on_integration_points(self, u, nb_degree_of_freedom=None, elements=None):
        u = np.asarray(u, dtype=np.float64)
        nb = nb_degree_of_freedom or (u.shape[1] if u.ndim == 2 else 1)
        return self._op(lib.akx_fe_interpolate_on_integration_points, u, nb, nb, elements)

    def compute_btd(self, Ds, elements=None):
        Ds = np.asarray(Ds, dtype=np.float64)
        rows = Ds.shape[1] // self.dim
        return self._op(lib.akx_fe_compute_btd, Ds, rows, rows * self.nb_nodes_per_element,
                        elements)

    def compute_ntb(self, bs, elements=None):
        bs = np.asarray(bs, dtype=np.float64)
        nb = bs.shape[1] if bs.ndim == 2 else 1
        return self._op(lib.akx_fe_compute_ntb, bs, nb, nb * self.nb_nodes_per_element, elements)

    def integrate(self, f, nb_degree_of_freedom=None, elements=None):
        """Per-element integrals (n_el, nb_dof) of a per-point field."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        nb = nb_degree_of_freedom or (f.shape[1] if f.ndim == 2 else 1)
        p, n, keep = _elements(elements)
        out = np.zeros((self._count(elements), nb))
        self.ctx.check(lib.akx_fe_integrate(self.ctx.h, _ptr(f), int(nb), p, n, _ptr(out)))
        return out

    def integrate_scalar(self, f, elements=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        p, n, keep = _elements(elements)
        v = ctypes.c_double()
        self.ctx.check(lib.akx_fe_integrate_scalar(self.ctx.h, _ptr(f), p, n, ctypes.byref(v)))
        return v.value

    def integrate_on_integration_points(self, f, nb_degree_of_freedom=None, elements=None):
        f = np.asarray(f, dtype=np.float64)
        nb = nb_degree_of_freedom or (f.shape[1] if f.ndim == 2 else 1)
        return self._op(lib.akx_fe_integrate_on_integration_points, f, nb, nb, elements)

    def assemble_field_lumped(self, field, lumped=None):
        """lumped (N, nb_dof) += row-sum lumping of a per-point field."""
        field = np.ascontiguousarray(field, dtype=np.float64)
        nb = field.shape[1] if field.ndim == 2 else 1
        if lumped is None:
            lumped = np.zeros((self.model.mesh.nb_nodes, nb))
        out = np.ascontiguousarray(lumped, dtype=np.float64)
        self.ctx.check(lib.akx_fe_assemble_field_lumped(self.ctx.h, _ptr(field), int(nb),
                                                        _ptr(out)))
        return out


class SolidMechanicsModel:
    """Explicit lumped-mass SolidMechanicsModel on one GPU (see module doc)."""

    def __init__(self, mesh, device: int = 0, keep_numbering: bool = False):
        self.mesh = mesh
        self.dim = mesh.dim
        self.ctx = Context(device, keep_numbering)
        self._X = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, dtype=np.uint32)
        self.ctx.check(lib.akx_mesh_set(self.ctx.h, mesh.elem_type, mesh.dim, mesh.nb_nodes,
                                        _ptr(self._X), mesh.nb_elements, _ptr(self._conn)))
        self.materials = []
        self.time_step = 0.0
        self._inflight = {}  # ticket -> PinnedBuffer of an asynchronous download

    # -- materials (MaterialFactory / input file `material elastic [...]`) --
    def add_material(self, kind=ELASTIC, E=None, nu=0.0, rho=None, plane_stress=False,
                     Sd=None, Yd=None):
        mid = ctypes.c_int(-1)
        self.ctx.check(lib.akx_material_add(self.ctx.h, kind, float(E), float(nu), float(rho),
                                            int(bool(plane_stress)), ctypes.byref(mid)))
        if kind == NEOHOOKEAN_DAMAGE:
            yd = np.ascontiguousarray(Yd, dtype=np.float64)
            self.ctx.check(lib.akx_material_damage(self.ctx.h, mid.value, float(Sd), _ptr(yd)))
        self.materials.append(dict(kind=kind, E=E, nu=nu, rho=rho))
        return mid.value

    def assign_materials(self, element_material):
        em = np.ascontiguousarray(element_material, dtype=np.int32)
        self.ctx.check(lib.akx_assign_materials(self.ctx.h, _ptr(em)))

    def init_full(self):
        self.ctx.check(lib.akx_init(self.ctx.h))

    # -- time stepping --
    def get_stable_time_step(self):
        dt = ctypes.c_double()
        self.ctx.check(lib.akx_stable_time_step(self.ctx.h, ctypes.byref(dt)))
        return dt.value

    def set_time_step(self, dt):
        self.time_step = float(dt)
        self.ctx.check(lib.akx_set_time_step(self.ctx.h, float(dt)))

    def solve_step(self, nsteps=1, precision=FP64):
        self.ctx.check(lib.akx_solve_step(self.ctx.h, int(nsteps), int(precision)))

    def solve_step_graph(self, nsteps=1, precision=FP64):
        """solve_step as one captured HIP graph launch (same results)."""
        self.ctx.check(lib.akx_solve_step_graph(self.ctx.h, int(nsteps), int(precision)))

    def set_node_time_steps(self, dt_per_node):
        """One time step per node (independent bodies in one model); None resets."""
        if dt_per_node is None:
            self.ctx.check(lib.akx_set_node_time_steps(self.ctx.h, None))
            return
        dt = np.ascontiguousarray(dt_per_node, dtype=np.float64)
        if dt.shape != (self.mesh.nb_nodes,):
            raise ValueError("one time step per node")
        self.ctx.check(lib.akx_set_node_time_steps(self.ctx.h, _ptr(dt)))

    def get_stable_time_step_per_material(self):
        out = np.zeros(len(self.materials), dtype=np.float64)
        self.ctx.check(lib.akx_stable_time_step_per_material(self.ctx.h, _ptr(out)))
        return out

    def assemble_internal_forces(self, precision=FP64):
        self.ctx.check(lib.akx_assemble_internal_forces(self.ctx.h, int(precision)))

    def get_energy(self, kind):
        v = ctypes.c_double()
        self.ctx.check(lib.akx_energy(self.ctx.h, ENERGIES[kind], ctypes.byref(v)))
        return v.value

    def set_dirichlet_increment(self, nodes, axis, increment):
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        self.ctx.check(lib.akx_set_dirichlet_increment(self.ctx.h, nodes.size, _ptr(nodes),
                                                       int(axis), float(increment)))

    def synchronize(self):
        self.ctx.check(lib.akx_synchronize(self.ctx.h))

    def enable_timing(self, on=True):
        self.ctx.check(lib.akx_enable_timing(self.ctx.h, int(bool(on))))

    def last_step_timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(lib.akx_last_step_timing(self.ctx.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_launch(self, element_block=0, node_block=0, xcd_remap=-1, mixed_recompute=-1,
                   slot_layout=-1, node_stage=-1):
        self.ctx.check(lib.akx_set_launch(self.ctx.h, element_block, node_block, xcd_remap,
                                          mixed_recompute, slot_layout, node_stage))

    def query(self, key):
        v = ctypes.c_double()
        self.ctx.check(lib.akx_query(self.ctx.h, key.encode(), ctypes.byref(v)))
        return v.value

    LAUNCH_KEYS = ("element_block", "node_block", "xcd_remap", "mixed_recompute",
                   "slot_layout", "node_stage")

    def autotune(self, precision=MIXED, nsteps=10):
        """Pick the fastest launch shape on this mesh; returns (shape, ms/step)."""
        chosen = np.zeros(6, dtype=np.int32)
        best = ctypes.c_double()
        self.ctx.check(lib.akx_autotune(self.ctx.h, int(precision), int(nsteps), _ptr(chosen),
                                        ctypes.byref(best)))
        shape = {k: int(v) for k, v in zip(self.LAUNCH_KEYS, chosen)}
        return shape, best.value

    # -- FEEngine and Material operators (fe_engine.hh, material.hh:110-195) --
    def get_fe_engine(self):
        return FEEngine(self)

    def material_elements(self, mat_id):
        n = ctypes.c_int64()
        self.ctx.check(lib.akx_material_elements(self.ctx.h, int(mat_id), ctypes.byref(n), None))
        ids = np.zeros(n.value, dtype=np.uint32)
        self.ctx.check(lib.akx_material_elements(self.ctx.h, int(mat_id), ctypes.byref(n),
                                                 _ptr(ids)))
        return ids

    def material_compute_stress(self, mat_id, grad_u):
        """Material::computeStress on the material's points from given gradients."""
        g = np.ascontiguousarray(grad_u, dtype=np.float64)
        out = np.zeros_like(g)
        self.ctx.check(lib.akx_material_compute_stress(self.ctx.h, int(mat_id), _ptr(g),
                                                       _ptr(out)))
        return out

    def material_compute_all_stresses(self, mat_id):
        """Material::computeAllStresses; returns (grad_u, stress) of its points."""
        rows = len(self.material_elements(mat_id)) * self.mesh.nb_qp
        g = np.zeros((rows, self.dim * self.dim))
        s = np.zeros_like(g)
        self.ctx.check(lib.akx_material_compute_all_stresses(self.ctx.h, int(mat_id), _ptr(g),
                                                             _ptr(s)))
        return g, s

    def material_assemble_internal_forces(self, mat_id):
        """Material::assembleInternalForces: adds into the internal force."""
        self.ctx.check(lib.akx_material_assemble_internal_forces(self.ctx.h, int(mat_id)))

    def material_potential_energy(self, mat_id):
        v = ctypes.c_double()
        self.ctx.check(lib.akx_material_potential_energy(self.ctx.h, int(mat_id),
                                                         ctypes.byref(v)))
        return v.value

    # -- constitutive surrogate (neohookean_damage on tetrahedron_4) --
    def surrogate_setup(self, mat_id, eps_max):
        self.ctx.check(lib.akx_surrogate_setup(self.ctx.h, int(mat_id), float(eps_max)))

    def surrogate_generate(self, paths, steps, inc_std=2e-4, yd_scale=1e6, yd_shape=5.0,
                           seed=1):
        self.ctx.check(lib.akx_surrogate_generate(self.ctx.h, int(paths), int(steps),
                                                  float(inc_std), float(yd_scale),
                                                  float(yd_shape), int(seed)))

    def surrogate_train(self, iterations, batch=8192, lr=2e-3, seed=7):
        loss = ctypes.c_double()
        self.ctx.check(lib.akx_surrogate_train(self.ctx.h, int(iterations), int(batch),
                                               float(lr), int(seed), ctypes.byref(loss)))
        return loss.value

    def surrogate_params(self, values=None):
        """Download (values None) or upload the FP32 parameter vector."""
        n = int(self.query("surrogate_parameters"))
        if values is None:
            out = np.zeros(n, dtype=np.float32)
            self.ctx.check(lib.akx_surrogate_params(self.ctx.h, _ptr(out), n, 0))
            return out
        arr = np.ascontiguousarray(values, dtype=np.float32)
        self.ctx.check(lib.akx_surrogate_params(self.ctx.h, _ptr(arr), arr.size, 1))

    def surrogate_gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(int(self.query("surrogate_parameters")), dtype=np.float32)
        loss = ctypes.c_double()
        self.ctx.check(lib.akx_surrogate_gradients(self.ctx.h, x.shape[0], _ptr(x), _ptr(t),
                                                   _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def surrogate_infer(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        self.ctx.check(lib.akx_surrogate_infer(self.ctx.h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def surrogate_dataset(self, first, count):
        x = np.zeros((count, 7), dtype=np.float32)
        t = np.zeros((count, 7), dtype=np.float32)
        self.ctx.check(lib.akx_surrogate_dataset(self.ctx.h, int(first), int(count), _ptr(x),
                                                 _ptr(t)))
        return x, t

    def surrogate_enable(self, on=True):
        self.ctx.check(lib.akx_surrogate_enable(self.ctx.h, int(bool(on))))

    def surrogate_fallbacks(self, reset=False):
        v = ctypes.c_uint64()
        self.ctx.check(lib.akx_surrogate_fallbacks(self.ctx.h, ctypes.byref(v), int(bool(reset))))
        return v.value

    def surrogate_format(self, fmt=None):
        """16-bit format of the inference kernel, "fp16" (default) or "bf16";
        with no argument, the current one."""
        if fmt is not None:
            self.ctx.check(lib.akx_surrogate_format(self.ctx.h, SURROGATE_FORMATS[fmt]))
        code = int(self.query("surrogate_format"))
        return {v: k for k, v in SURROGATE_FORMATS.items()}[code]

    # -- output (SolidMechanicsModel::setBaseName / addDumpField / dump) --
    def set_base_name(self, name, directory="paraview", asynchronous=True, binary=True):
        from .dumper import Dumper
        if getattr(self, "_dumper", None) is not None:
            self._dumper.close()
        self._dumper = Dumper(self, name, directory, asynchronous, binary=binary)

    def add_dump_field(self, name):
        if getattr(self, "_dumper", None) is None:
            self.set_base_name("model")
        self._dumper.add_dump_field(name)

    def dump(self, time=None):
        if getattr(self, "_dumper", None) is None:
            self.set_base_name("model")
        return self._dumper.dump(time)

    # -- fields --
    def _shape(self, name):
        m, d = self.mesh, self.dim
        eq = m.nb_elements * m.nb_qp
        if name in ("grad_u", "stress"):
            return (eq, d * d)
        if name in ("damage", "jacobians"):
            return (eq,)
        if name == "shapes_derivatives":
            return (eq, d * m.conn.shape[1])
        return (m.nb_nodes, d)

    def get(self, name):
        dtype = np.uint8 if name == "blocked_dofs" else np.float64
        out = np.zeros(self._shape(name), dtype=dtype)
        self.ctx.check(lib.akx_field_download(self.ctx.h, FIELDS[name], _ptr(out), out.nbytes))
        if name == "blocked_dofs":
            return out.astype(bool)
        return out

    def field_layout(self, name):
        """(shape, dtype) of a field as get() returns it before any conversion."""
        return self._shape(name), (np.uint8 if name == "blocked_dofs" else np.float64)

    def get_async(self, name, buf):
        """Snapshot `name` now; its bytes arrive in `buf` (a PinnedBuffer) while
        later steps run.  Returns a ticket for wait()."""
        shape, dtype = self.field_layout(name)
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if buf.nbytes < nbytes:
            raise ValueError("pinned buffer too small for %s" % name)
        t = ctypes.c_uint64()
        self.ctx.check(lib.akx_field_download_async(self.ctx.h, FIELDS[name], buf.ptr, nbytes,
                                                    ctypes.byref(t)))
        # the buffer stays referenced until its copy is known to be done, so
        # that dropping it early cannot free memory the copy engine writes
        self._inflight[t.value] = buf
        return t.value

    def wait(self, ticket):
        """Block until the asynchronous download `ticket` has landed (may be
        called from another thread than the one that steps the model)."""
        self.ctx.check(lib.akx_download_wait(self.ctx.h, ticket))
        self._inflight.pop(ticket, None)

    def done(self, ticket):
        flag = ctypes.c_int()
        self.ctx.check(lib.akx_download_query(self.ctx.h, ticket, ctypes.byref(flag)))
        if flag.value:
            self._inflight.pop(ticket, None)
        return bool(flag.value)

    def set(self, name, values):
        dtype = np.uint8 if name == "blocked_dofs" else np.float64
        arr = np.ascontiguousarray(np.broadcast_to(np.asarray(values).astype(dtype),
                                                   self._shape(name)))
        self.ctx.check(lib.akx_field_upload(self.ctx.h, FIELDS[name], _ptr(arr), arr.nbytes))

    # reference-style accessors
    def get_displacement(self):
        return self.get("displacement")

    def get_velocity(self):
        return self.get("velocity")

    def get_acceleration(self):
        return self.get("acceleration")

    def get_internal_force(self):
        return self.get("internal_force")

    def get_mass(self):
        return self.get("mass")

    def get_blocked_dofs(self):
        return self.get("blocked_dofs")

    def flush_dumps(self):
        """Wait until every dump so far is written."""
        if getattr(self, "_dumper", None) is not None:
            self._dumper.flush()

    def close(self):
        if getattr(self, "_dumper", None) is not None:
            self._dumper.close()
            self._dumper = None
        self.ctx.close()
