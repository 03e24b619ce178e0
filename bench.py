#!/usr/bin/env python3
"""Benchmark of the explicit central-difference step on one MI355X.

Default workload (cube-tet4): 3D linear elastic unit cube, 100^3 hexes
Kuhn-split into 6,000,000 tetrahedron_4 (1,030,301 nodes), interior nodes
jittered by +-1 % of h, E = 210 GPa, nu = 0.3, rho = 7800 kg/m^3, bottom face
blocked, top face v_z = -1 m/s, dt = 0.8 x stable.  A "step" is one
solveStep(): predictor, internal forces (gradient -> stress -> B^T sigma ->
integrate -> assembly), residual, lumped solve, corrector.

value = element-QP updates per second of the mixed path (FP32 element kernels,
FP64 nodal state), inputs resident in HBM, timed over --steps steps between
two stream synchronisations.  The FP64 path is timed the same way and reported
beside it, with the relative nodal-displacement error of the mixed path after
1000 steps.  The CPU baseline is the oracle (C++ FP64 restatement of the
reference loops) timed on this host over a bounded number of steps.

Other workloads (--workload), reported the same way on their own JSON line:
  hex8-neohookean  128^3 hexahedron_8 (2x2x2 Gauss), finite-strain neo-Hookean
                   E = 10 MPa, nu = 0.45, rho = 1000, bottom blocked, top face
                   compressed at 0.5 m/s (per-step Dirichlet increment), 500 steps
                   for the error run;
  ensemble         512 independent plane-strain triangle_3 specimens (2,000 to
                   40,000 elements each), each with its own material and stable
                   time step, advanced together by one captured HIP graph per
                   run; 2000 steps for the per-specimen error run;
  plate-surrogate  notched plate, ~4M tetrahedron_4, neo-Hookean + damage: the
                   plain law against the MLP surrogate (trained on the GPU
                   first), with the surrogate's stress error on the run's own
                   strain states.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "element-QP updates/s per explicit step, mixed vs FP64 at ≤1e-5 rel. error"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(E, N, n, Q, d, pe, has_fext=False, tables=True, node_ordered=False,
                      patch_nodes=None, patch_blocks=0, dt_per_node=False, state_bytes=0,
                      partials=0, packed=False, partials_node_major=False, position=False):
    """Compulsory HBM bytes of one step, each array touched once.
    pe: bytes of the element tables and force slots (4 mixed, 8 FP64);
    tables=False: B and det(J) w rebuilt from the coordinates (one more nodal
    gather instead of the two tables); node_ordered: the CSR slot positions
    are read by the element kernel (scatter) instead of the node kernel;
    patch_nodes: LDS-patch element kernel, which reads 16-bit local
    connectivity + 8-bit run ranks per (e, a) and, per patch node, its id,
    run start and run offset (4 + 4 + 2 bytes); dt_per_node: one time step
    per node (ensembles); state_bytes: per-point history read + written;
    partials: bytes per value of the one sum per patch node that the patch
    kernel writes (8 FP64, 4 FP32; no run start needed) and the node kernel
    gathers through a node -> entry CSR instead of reading E*n slots;
    packed: local node and run rank in one 16-bit word per (e, a), 3-node
    elements padded to 4; partials_node_major: the patch kernel reads each
    partial's node-major position and the node kernel streams its partials
    (no entry list); position: the node kernel steps in the position form
    (mixed path) and reads u and the previous displacement and writes u,
    neither v nor a, and dt^2 / M per node in place of the mass and the
    per-node time step."""
    state = (3 if position else 6) * N * d * 8
    local = (E * (4 if n == 3 else n) * 2) if packed else E * n * (2 + 1)
    if partials:
        pos = patch_nodes * 4
        elem = (local + patch_nodes * (4 + 2) + patch_blocks * (4 + 2)
                + 2 * N * d * 8 + patch_nodes * d * partials + state_bytes
                + (pos if partials_node_major else 0))
        node = ((N + 1) * 4 + (0 if partials_node_major else pos) + patch_nodes * d * partials
                + state + N * 8 + N * d
                + (N * d * 8 if has_fext else 0)
                + (N * 8 if dt_per_node and not position else 0))
        return elem, node
    if patch_nodes is not None:
        elem = (local                                      # local connectivity, run ranks
                + patch_nodes * (4 + 4 + 2) + patch_blocks * (4 + 2)  # patch node lists
                + 2 * N * d * 8                            # u and X of every node
                + E * n * d * pe)                          # force slots (write)
        node_ordered = True
    else:
        elem = (E * n * 4                                   # connectivity
                + (E * Q * d * n * pe if tables else 0)      # B = dN/dX
                + (E * Q * pe if tables else N * d * 8)      # det(J) w, or X gather
                + N * d * 8                                  # displacement gather
                + (E * n * 4 if node_ordered else 0)         # slot positions
                + E * n * d * pe)                            # element force slots (write)
    elem += state_bytes
    node = ((N + 1) * 4 + (0 if node_ordered else E * n * 4)  # node -> slot CSR
            + E * n * d * pe         # force slots (read)
            + state                  # u, v (and a) read + write
            + N * 8                  # lumped mass (one value per node)
            + N * d                  # blocked flags
            + (N * d * 8 if has_fext else 0)
            + (N * 8 if dt_per_node and not position else 0))
    return elem, node


def fused_bytes(E, N, n, d, evaluations, nblk, node_block=256, chunk=256):
    """Compulsory HBM bytes of one fused FP64 step (slot layout 6): the blocks'
    element lists (4 bytes per evaluation) and LDS slots (2 bytes per node of
    an evaluation), the per-chunk run tables, connectivity and coordinates
    once, u read and written (two buffers), v and a read and written, lumped
    mass and blocked flags."""
    chunks = evaluations / chunk + nblk
    return (evaluations * (4 + 2 * n) + chunks * (node_block + 1) * 2 + (nblk + 1) * 8
            + E * n * 4 + N * d * 8 + 2 * N * d * 8 + 4 * N * d * 8 + N * 8 + N * d)


def fused_mixed_bytes(E, N, n, d, evaluations, nblk, dt_per_node=False, node_block=256,
                      chunk=256, records=True):
    """Compulsory HBM bytes of one fused mixed step (slot layout 6 on the mixed
    path): as the FP64 one, but within a run the node update is the position
    form (the previous displacement read from the output buffer before it is
    overwritten; neither v nor a; dt^2 / M per node in place of the mass and
    the per-node time step), and triangles with
    records (mixed_recompute 1) stream one 32-byte record per evaluation
    (coordinate differences, node ids, LDS slots, material) instead of the
    element list, LDS slots, connectivity and coordinates."""
    chunks = evaluations / chunk + nblk
    tables = chunks * (node_block + 1) * 2 + (nblk + 1) * 8
    if n == 3 and records:
        elements = evaluations * 32
    else:
        elements = evaluations * (4 + 2 * n) + E * n * 4 + N * d * 8
    return elements + tables + 3 * N * d * 8 + N * 8 + N * d


def make_workload(n):
    from akantu_amd import mesh as akm
    m = akm.cube_tet4(n, 0.01, 1)
    X = m.nodes
    blocked = np.zeros_like(X, dtype=bool)
    blocked[np.isclose(X[:, 2], 0.0)] = True
    v0 = np.zeros_like(X)
    v0[np.isclose(X[:, 2], 1.0), 2] = -1.0
    return m, blocked, v0


def reset(model, v0):
    z = np.zeros_like(v0)
    model.set("displacement", z)
    model.set("velocity", v0)
    model.set("acceleration", z)


def run_steps(model, precision, steps, graph):
    if graph:
        model.solve_step_graph(steps, precision)
    else:
        model.solve_step(steps, precision)


def timed(model, precision, steps, warmup, v0, graph=False):
    reset(model, v0)
    if warmup:
        run_steps(model, precision, warmup, graph)
    model.synchronize()
    t0 = time.perf_counter()
    run_steps(model, precision, steps, graph)
    model.synchronize()
    return time.perf_counter() - t0


def kernel_times(model, precision, steps, v0):
    """Average element / node kernel durations from HIP events recorded on the
    model's stream around every launch."""
    reset(model, v0)
    model.enable_timing(True)
    model.solve_step(steps, precision)
    model.synchronize()
    el, nd = model.last_step_timing()
    model.enable_timing(False)
    return el, nd


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Device-to-device hipMemcpy of `nbytes` (the runtime's copy kernel):
    the measured streaming ceiling, 2 * nbytes moved per copy."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        return None
    try:
        hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
        for _ in range(2):
            hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)
        hip.hipDeviceSynchronize()
        t = (time.perf_counter() - t0) / reps
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return 2 * nbytes / t / 1e9


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(m, materials, blocked, v0, dt, warm=2, steps=10, threads=1):
    """The oracle on this host: `steps` timed steps after `warm` untimed ones."""
    from oracle import oracle as orc
    o = orc.OracleModel(m, materials, threads=threads)
    o.blocked[:] = blocked
    o.nodal("v")[:] = v0
    o.set_time_step(dt)
    o.solve_step(warm)
    t0 = time.perf_counter()
    o.solve_step(steps)
    t = time.perf_counter() - t0
    return o, t / steps


def tune(model, args, akx, profiled=None):
    """Launch shapes: given on the command line, else the ones the committed
    profiles were taken with (`profiled`, a traffic JSON under profiles/,
    written from an earlier autotuned run), else the autotuner's choice."""
    shapes = {}
    keys = akx.SolidMechanicsModel.LAUNCH_KEYS
    stored = {}
    if profiled and not args.autotune:
        path = os.path.join(ROOT, "profiles", profiled)
        if os.path.exists(path):
            with open(path) as f:
                tr = json.load(f)
            for name in ("mixed", "fp64"):
                sh = tr.get("launch_" + name)
                if sh and set(sh) == set(keys):
                    stored[name] = sh
    for name, prec, fixed in (("mixed", akx.MIXED, args.launch_mixed),
                              ("fp64", akx.FP64, args.launch_fp64)):
        if not fixed and name in stored:
            shapes[name] = dict(stored[name])
            log("%s: launch of the committed profiles %s" % (name, shapes[name]))
        elif fixed:
            shapes[name] = dict(zip(keys, (int(x) for x in fixed.split(","))))
            log("%s: fixed launch %s" % (name, shapes[name]))
        else:
            shapes[name], tuned_ms = model.autotune(prec, 10)
            log("%s: autotuned launch %s (%.4f ms/step)" % (name, shapes[name], tuned_ms))
    return shapes


def measure(model, shapes, args, akx, v0, QP, graph=False):
    res = {}
    for name, prec in (("mixed", akx.MIXED), ("fp64", akx.FP64)):
        model.set_launch(**shapes[name])
        t = timed(model, prec, args.steps, args.warmup, v0, graph)
        el_ms, nd_ms = kernel_times(model, prec, min(args.steps, 200), v0)
        res[name] = dict(seconds=t, ms_per_step=1e3 * t / args.steps,
                         qp_per_s=QP * args.steps / t, element_ms=el_ms, node_ms=nd_ms,
                         launch=shapes[name])
        log("%s: %.4f ms/step  (element %.4f ms, node %.4f ms)" %
            (name, res[name]["ms_per_step"], el_ms, nd_ms))
    return res


def dominant(res, shapes, model, elem_kernel, node_kernel, E, N, n, Q, d, **kw):
    """Roofline of the slower of the two mixed-path kernels."""
    lay = shapes["mixed"]["slot_layout"]
    patch = lay >= 2 and model.query("patch_available") == 1
    partials = {3: 8, 4: 4, 5: 4}.get(lay, 0) if patch else 0
    eb, nb = algorithmic_bytes(E, N, n, Q, d, 4,
                               tables=not (shapes["mixed"]["mixed_recompute"] or patch),
                               node_ordered=bool(lay),
                               patch_nodes=model.query("patch_nodes_total") if patch else None,
                               patch_blocks=model.query("patch_blocks") if patch else 0,
                               partials=partials,
                               packed=patch and model.query("patch_packed") == 1,
                               partials_node_major=patch and lay == 5, position=True, **kw)
    el_ms, nd_ms = res["mixed"]["element_ms"], res["mixed"]["node_ms"]
    if el_ms >= nd_ms:
        kname, kbytes, kms = elem_kernel(patch, shapes["mixed"]["mixed_recompute"]), eb, el_ms
    else:
        # per-patch partials: the node kernel's partial-sum instance
        kname = (node_kernel.replace(", 1, ", ", %d, " % (2 if partials == 8 else 3))
                 if partials else node_kernel)
        kbytes, kms = nb, nd_ms
    achieved = kbytes / (kms * 1e-3) / 1e9
    step_gbs = (eb + nb) / (res["mixed"]["ms_per_step"] * 1e-3) / 1e9
    return kname, kbytes, kms, achieved, eb, nb, step_gbs


def base_record(args, res, dtype, config):
    return {
        "metric": METRIC,
        "value": res["mixed"]["qp_per_s"],
        "unit": "QP-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
        "mixed_over_fp64": res["mixed"]["qp_per_s"] / res["fp64"]["qp_per_s"],
    }


def traffic_for(kname, shapes, profile_json):
    """HBM bytes per dispatch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs, valid for the launch shape they were taken with."""
    tpath = os.path.join(ROOT, "profiles", profile_json)
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        tr = json.load(f)
    k = tr.get("kernels", {}).get(kname)
    if k and tr.get("launch_mixed") == shapes["mixed"]:
        return k["fetch_bytes"] + k["write_bytes"]
    return None


def fp64_record(res, shapes, step_bytes=None):
    r = res["fp64"]
    out = {"value": r["qp_per_s"], "ms_per_step": r["ms_per_step"],
           "element_ms": r["element_ms"], "node_ms": r["node_ms"], "launch": shapes["fp64"]}
    if step_bytes:
        out["step_gbs"] = step_bytes / (r["ms_per_step"] * 1e-3) / 1e9
    return out


# ---------------------------------------------------------------------------
def bench_cube(args, akx):
    log("building workload: cube %d^3 hexes -> tetrahedron_4" % args.n)
    m, blocked, v0 = make_workload(args.n)
    E, N, n, Q, d = m.nb_elements, m.nb_nodes, 4, 1, 3
    log("E=%d N=%d" % (E, N))
    mat = dict(E=210e9, nu=0.3, rho=7800.0)
    model = akx.SolidMechanicsModel(m)
    model.add_material(**mat)
    model.init_full()
    model.set("blocked_dofs", blocked)
    reset(model, v0)
    dt = 0.8 * model.get_stable_time_step()
    model.set_time_step(dt)
    log("dt=%.6e" % dt)

    shapes = tune(model, args, akx, "cube_tet4_traffic.json" if args.n == 100 else None)
    res = measure(model, shapes, args, akx, v0, E * Q)

    # accuracy of the mixed path against the FP64 path after error-steps steps
    log("error run: %d steps mixed and FP64" % args.error_steps)
    reset(model, v0)
    model.set_launch(**shapes["mixed"])
    model.solve_step(args.error_steps, akx.MIXED)
    u_mixed = model.get_displacement()
    e_mixed = model.get_energy("kinetic") + model.get_energy("potential")
    reset(model, v0)
    model.set_launch(**shapes["fp64"])
    model.solve_step(args.error_steps, akx.FP64)
    u_fp64 = model.get_displacement()
    e_fp64 = model.get_energy("kinetic") + model.get_energy("potential")
    rel_u = float(np.linalg.norm(u_mixed - u_fp64) / np.linalg.norm(u_fp64))
    rel_e = float(abs(e_mixed - e_fp64) / abs(e_fp64))
    log("mixed vs FP64 after %d steps: rel u error %.3e, energy %.3e" %
        (args.error_steps, rel_u, rel_e))

    def elem_kernel(patch, rc):
        return ("void akx::k_element_patch<2, 0>(akx::ElemArgs)" if patch else
                "void akx::k_element_mixed<2, 0, %s>(akx::ElemArgs)" % ("true" if rc else "false"))
    kname, kbytes, kms, achieved, eb, nb, step_gbs = dominant(
        res, shapes, model, elem_kernel, "void akx::k_node<3, 1, false>(akx::NodeArgs)",
        E, N, n, Q, d)
    traffic = traffic_for(kname, shapes, "cube_tet4_traffic.json") if args.n == 100 else None
    eb64, nb64 = algorithmic_bytes(E, N, n, Q, d, 8,
                                   tables=not shapes["fp64"]["mixed_recompute"],
                                   node_ordered=bool(shapes["fp64"]["slot_layout"]))
    if shapes["fp64"]["slot_layout"] == 6 and model.query("fused_available") == 1:
        eb64 = fused_bytes(E, N, n, d, model.query("fused_elements"), (N + 255) // 256)
        nb64 = 0
    copy_gbs = copy_bandwidth()
    log("device copy ceiling: %s GB/s" % (None if copy_gbs is None else round(copy_gbs, 1)))

    cpu = None
    cpu_all = None
    bitwise = None
    if not args.no_cpu:
        threads = 1
        log("CPU baseline: oracle, %d thread, %d steps (+2 warm-up) on the full mesh"
            % (threads, args.cpu_steps))
        o, sec = cpu_baseline(m, [mat], blocked, v0, dt, 2, args.cpu_steps, threads)
        cpu = {"value": E * Q / sec, "unit": "QP-updates/s", "cores": threads, "kind": "port",
               "sample": "oracle (C++ FP64 restatement of the reference loops, g++ -O3 "
                         "-ffp-contract=off), full %d-element mesh, %d timed steps after 2 "
                         "warm-up steps, %.3f s/step" % (E, args.cpu_steps, sec)}
        log("CPU: %.3f s/step" % sec)
        # the same on all the host cores this job may use (the oracle's element
        # partition with per-thread force arrays reduced in a fixed order)
        nthr = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))
        _, sec_all = cpu_baseline(m, [mat], blocked, v0, dt, 2, args.cpu_steps, nthr)
        cpu_model = host_cpu_model()
        host_cpus = os.cpu_count() or 0
        cpu_all = {"value": E * Q / sec_all, "unit": "QP-updates/s", "cores": nthr,
                   "host_cpus": host_cpus, "kind": "port",
                   "sample": "oracle on %d threads (the job's CPU share; the host shows %d "
                             "CPUs), %s, full %d-element mesh, %d timed steps after 2 warm-up "
                             "steps, %.3f s/step; per-thread force arrays allocated once, "
                             "reduced in parallel over DOF ranges in thread order"
                             % (nthr, host_cpus, cpu_model, E, args.cpu_steps, sec_all)}
        cpu["sample"] += ", " + cpu_model
        log("CPU, %d threads: %.3f s/step" % (nthr, sec_all))
        # the FP64 GPU path against the oracle on the full mesh, same step count
        reset(model, v0)
        model.set_launch(**shapes["fp64"])
        model.solve_step(2 + args.cpu_steps, akx.FP64)
        bitwise = bool(np.array_equal(model.get_displacement(), o.nodal("u")) and
                       np.array_equal(model.get_velocity(), o.nodal("v")))
        log("FP64 GPU == oracle after %d steps on the full mesh: %s" % (2 + args.cpu_steps, bitwise))
        del o

    out = base_record(args, res, "f32 element kernels / f64 nodal state", {
        "workload": "cube-tet4: 3D linear elastic cube, %d^3 hexes Kuhn-split into %d "
                    "tetrahedron_4 (1 QP each), %d nodes, +-1%% interior jitter, "
                    "E=210 GPa nu=0.3 rho=7800, bottom blocked, top v_z=-1 m/s, "
                    "dt=0.8 x stable" % (args.n, E, N),
        "elements": E, "nodes": N, "qp_per_element": Q, "precision": "mixed",
        "parallelism": "1 GPU"})
    out.update({
        "fp64": fp64_record(res, shapes, eb64 + nb64),
        "launch_mixed": shapes["mixed"],
        "accuracy": {"steps": args.error_steps, "rel_u_error_mixed_vs_fp64": rel_u,
                     "rel_energy_error_mixed_vs_fp64": rel_e,
                     "fp64_gpu_bitwise_equal_oracle": bitwise},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": kbytes, "avg_ms": kms},
        "roofline_step": {"algorithmic_bytes": eb + nb, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "element_ms": res["mixed"]["element_ms"],
                          "node_ms": res["mixed"]["node_ms"],
                          "element_gbs": eb / (res["mixed"]["element_ms"] * 1e-3) / 1e9,
                          "node_gbs": nb / (res["mixed"]["node_ms"] * 1e-3) / 1e9},
        "copy_ceiling_gbs": copy_gbs,
        "cpu_baseline": cpu,
        "cpu_all_cores": cpu_all,
    })
    return out


def bench_hex8(args, akx):
    from akantu_amd import mesh as akm
    nh = args.n if args.n != 100 else 128
    log("building workload: cube %d^3 hexahedron_8, neo-Hookean" % nh)
    m = akm.cube_hex8(nh)
    E, N, n, Q, d = m.nb_elements, m.nb_nodes, 8, 8, 3
    X = m.nodes
    blocked = np.zeros_like(X, dtype=bool)
    blocked[np.isclose(X[:, 2], 0.0)] = True
    top = np.nonzero(np.isclose(X[:, 2], 1.0))[0]
    v0 = np.zeros_like(X)
    model = akx.SolidMechanicsModel(m)
    model.add_material(kind=akx.NEOHOOKEAN, E=1e7, nu=0.45, rho=1000.0)
    model.init_full()
    model.set("blocked_dofs", blocked)
    dt = 0.8 * model.get_stable_time_step()
    model.set_time_step(dt)
    model.set_dirichlet_increment(top, 2, -0.5 * dt)
    reset(model, v0)
    log("E=%d N=%d dt=%.6e" % (E, N, dt))
    profiled = "hex8_neohookean_traffic.json" if nh == 128 else None
    shapes = tune(model, args, akx, profiled)
    res = measure(model, shapes, args, akx, v0, E * Q)
    steps = args.error_steps if args.error_steps != 1000 else 500
    log("error run: %d steps mixed and FP64" % steps)
    reset(model, v0)
    model.set_launch(**shapes["mixed"])
    model.solve_step(steps, akx.MIXED)
    u_mixed = model.get_displacement()
    reset(model, v0)
    model.set_launch(**shapes["fp64"])
    model.solve_step(steps, akx.FP64)
    u_fp64 = model.get_displacement()
    rel_u = float(np.linalg.norm(u_mixed - u_fp64) / np.linalg.norm(u_fp64))
    log("mixed vs FP64 after %d steps: rel u error %.3e" % (steps, rel_u))
    strain = strain_run(model, shapes, akx, v0, top, dt) if not args.no_strain_run else None

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))
        cpu_steps = max(1, args.cpu_steps // 5)
        log("CPU baseline: oracle, %d threads, %d steps (+1 warm-up) on the full mesh"
            % (threads, cpu_steps))
        vt = v0.copy()
        _, sec = cpu_baseline(m, [dict(kind=akx.NEOHOOKEAN, E=1e7, nu=0.45, rho=1000.0)],
                              blocked, vt, dt, 1, cpu_steps, threads)
        cpu = {"value": E * Q / sec, "unit": "QP-updates/s", "cores": threads,
               "host_cpus": os.cpu_count() or 0, "kind": "port",
               "sample": "oracle (C++ FP64 restatement of the reference loops), %d threads, "
                         "%s, full %d-element mesh (%d points), %d timed steps after 1 warm-up "
                         "step, %.3f s/step" % (threads, host_cpu_model(), E, E * Q, cpu_steps,
                                                sec)}
        log("CPU: %.3f s/step" % sec)

    def elem_kernel(patch, rc):
        return ("void akx::k_element_patch<3, 1>(akx::ElemArgs)" if patch else
                "void akx::k_element_mixed<3, 1, %s>(akx::ElemArgs)" % ("true" if rc else "false"))
    kname, kbytes, kms, achieved, eb, nb, step_gbs = dominant(
        res, shapes, model, elem_kernel, "void akx::k_node<3, 1, false>(akx::NodeArgs)",
        E, N, n, Q, d)
    traffic = traffic_for(kname, shapes, profiled) if profiled else None
    out = base_record(args, res, "f32 element kernels / f64 nodal state", {
        "workload": "hex8-neohookean: %d^3 hexahedron_8 (2x2x2 Gauss), %d nodes, finite-strain "
                    "neo-Hookean E=10 MPa nu=0.45 rho=1000, bottom blocked, top compressed "
                    "at 0.5 m/s, dt=0.8 x stable" % (nh, N),
        "elements": E, "nodes": N, "qp_per_element": Q, "precision": "mixed",
        "parallelism": "1 GPU"})
    out.update({
        "fp64": fp64_record(res, shapes),
        "launch_mixed": shapes["mixed"],
        "accuracy": {"steps": steps, "rel_u_error_mixed_vs_fp64": rel_u,
                     "to_10_percent_strain": strain},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": kbytes, "avg_ms": kms},
        "roofline_step": {"algorithmic_bytes": eb + nb, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    })
    return out


def strain_run(model, shapes, akx, v0, top, dt, strain=0.10, speed=0.5):
    """The hexahedra compressed to `strain` (top face at `speed`), mixed and
    FP64: displacement error, and each path's energy balance: the change of
    E_kin + E_pot against the work of the moving face (getExternalWork summed
    every step; the face's nodes carry their prescribed velocity)."""
    nsteps = int(math.ceil(strain / (speed * dt)))
    log("strain run: %d steps to %.0f %% strain, mixed and FP64, energies every step"
        % (nsteps, 100 * strain))
    vt = v0.copy()
    vt[top, 2] = -speed
    out = {"steps": nsteps}
    u = {}
    for name, prec in (("mixed", akx.MIXED), ("fp64", akx.FP64)):
        reset(model, vt)
        model.set_launch(**shapes[name])
        e0 = model.get_energy("kinetic") + model.get_energy("potential")
        work = 0.0
        for _ in range(nsteps):
            model.solve_step(1, prec)
            work += model.get_energy("external work")
        e1 = model.get_energy("kinetic") + model.get_energy("potential")
        u[name] = model.get_displacement()
        out[name] = {"energy_change": e1 - e0, "boundary_work": work,
                     "balance_rel": (e1 - e0 - work) / work, "e_kin_plus_e_pot": e1}
    out["top_displacement"] = float(-u["fp64"][top, 2].mean())
    out["rel_u_error_mixed_vs_fp64"] = float(np.linalg.norm(u["mixed"] - u["fp64"]) /
                                             np.linalg.norm(u["fp64"]))
    out["rel_energy_error_mixed_vs_fp64"] = float(
        abs(out["mixed"]["e_kin_plus_e_pot"] - out["fp64"]["e_kin_plus_e_pot"]) /
        out["fp64"]["e_kin_plus_e_pot"])
    log("strain run: top face moved %.4f, rel u error %.3e, energy balance mixed %.2e fp64 %.2e"
        % (out["top_displacement"], out["rel_u_error_mixed_vs_fp64"],
           out["mixed"]["balance_rel"], out["fp64"]["balance_rel"]))
    return out


def bench_ensemble(args, akx):
    from akantu_amd import mesh as akm
    log("building workload: 512 triangle_3 specimens")
    en = akm.concatenate_specimens(akm.ensemble_tri3(512))
    m = en.mesh
    E, N, n, Q, d = m.nb_elements, m.nb_nodes, 3, 1, 2
    model = akx.SolidMechanicsModel(m)
    for s in en.specimens:
        model.add_material(E=s.E, nu=s.nu, rho=s.rho)
    model.assign_materials(en.element_material)
    model.init_full()
    model.set("blocked_dofs", en.blocked)
    v0 = en.velocity
    reset(model, v0)
    dt_mat = model.get_stable_time_step_per_material()
    model.set_node_time_steps(0.8 * dt_mat[en.node_specimen()])
    log("E=%d N=%d, specimen dt in [%.3e, %.3e]" % (E, N, 0.8 * dt_mat.min(), 0.8 * dt_mat.max()))
    shapes = tune(model, args, akx, "ensemble_tri3_traffic.json")
    res = measure(model, shapes, args, akx, v0, E * Q, graph=True)
    steps = args.error_steps if args.error_steps != 1000 else 2000
    chunks = 10
    log("error run: %d steps mixed and FP64 (graph), displacements sampled every %d steps"
        % (steps, steps // chunks))
    # Each specimen oscillates: its displacement norm at one instant can be
    # near zero, so the error is also reported against the largest norm the
    # specimen reaches over the run (sampled at the chunk ends).
    u = {}
    for name, prec in (("mixed", akx.MIXED), ("fp64", akx.FP64)):
        reset(model, v0)
        model.set_launch(**shapes[name])
        u[name] = []
        for _ in range(chunks):
            model.solve_step_graph(steps // chunks, prec)
            u[name].append(model.get_displacement())
    S = len(en.specimens)
    err_end, err_amp = np.zeros(S), np.zeros(S)
    for k in range(S):
        a, b = en.node_offsets[k], en.node_offsets[k + 1]
        amp = max(np.linalg.norm(uf[a:b]) for uf in u["fp64"])
        dev = max(np.linalg.norm(um[a:b] - uf[a:b]) for um, uf in zip(u["mixed"], u["fp64"]))
        err_amp[k] = dev / amp
        err_end[k] = (np.linalg.norm(u["mixed"][-1][a:b] - u["fp64"][-1][a:b]) /
                      np.linalg.norm(u["fp64"][-1][a:b]))
    log("per-specimen error vs amplitude: max %.3e median %.3e; at the last step: max %.3e "
        "median %.3e" % (err_amp.max(), np.median(err_amp), err_end.max(), np.median(err_end)))

    def elem_kernel(patch, rc):
        return ("void akx::k_element_patch<0, 0>(akx::ElemArgs)" if patch else
                "void akx::k_element_mixed<0, 0, %s>(akx::ElemArgs)" % ("true" if rc else "false"))
    fused = model.query("fused_available") == 1
    nblk = (N + 255) // 256
    step64 = None
    if fused and shapes["fp64"]["slot_layout"] == 6:
        step64 = fused_bytes(E, N, n, d, model.query("fused_elements"), nblk) + N * 8
    if fused and shapes["mixed"]["slot_layout"] == 6:
        # one kernel per step: its roofline is the step's
        records = (shapes["mixed"]["mixed_recompute"] == 1 and
                   model.query("fused_records") == 1)
        kname = ("void akx::k_step_fused_mixed<0, 0, false, %s>"
                 "(akx::ElemArgs, akx::NodeArgs, akx::FusedArgs)" % ("true" if records else "false"))
        kbytes = fused_mixed_bytes(E, N, n, d, model.query("fused_elements"), nblk,
                                   dt_per_node=True, records=records)
        kms = res["mixed"]["element_ms"]
        achieved = kbytes / (kms * 1e-3) / 1e9
        eb, nb = kbytes, 0
        step_gbs = kbytes / (res["mixed"]["ms_per_step"] * 1e-3) / 1e9
    else:
        kname, kbytes, kms, achieved, eb, nb, step_gbs = dominant(
            res, shapes, model, elem_kernel, "void akx::k_node<2, 1, false>(akx::NodeArgs)",
            E, N, n, Q, d, dt_per_node=True)
    traffic = traffic_for(kname, shapes, "ensemble_tri3_traffic.json")
    out = base_record(args, res, "f32 element kernels / f64 nodal state", {
        "workload": "ensemble: 512 plane-strain triangle_3 specimens, %d elements and %d nodes "
                    "in total, per-specimen E, nu, rho and stable dt, sinusoidal initial "
                    "velocity, one captured HIP graph per run" % (E, N),
        "elements": E, "nodes": N, "specimens": 512, "qp_per_element": Q,
        "precision": "mixed", "parallelism": "1 GPU"})
    out.update({
        "fp64": fp64_record(res, shapes, step64),
        "launch_mixed": shapes["mixed"],
        "accuracy": {"steps": steps,
                     "per_specimen_error_vs_amplitude_max": float(err_amp.max()),
                     "per_specimen_error_vs_amplitude_median": float(np.median(err_amp)),
                     "per_specimen_rel_u_error_last_step_max": float(err_end.max()),
                     "per_specimen_rel_u_error_last_step_median": float(np.median(err_end)),
                     "note": "error_vs_amplitude = max over the samples of |u_mixed - u_fp64| "
                             "/ max over the samples of |u_fp64| per specimen"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": kbytes, "avg_ms": kms},
        "roofline_step": {"algorithmic_bytes": eb + nb, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "cpu_baseline": None,
    })
    return out


def bench_plate_surrogate(args, akx):
    """Notched plate of tetrahedron_4 under tension, neo-Hookean + Marigo
    damage (Weibull thresholds): the plain law against the MLP surrogate in
    the mixed step, with the network trained on the GPU first."""
    from akantu_amd import mesh as akm
    nx = args.n if args.n != 100 else 150
    nz = max(2, nx // 5)
    log("building workload: notched plate %d x %d x %d cells -> tetrahedron_4" % (nx, nx, nz))
    m = akm.notched_plate_tet4(nx, nx, nz)
    E, N, n, Q, d = m.nb_elements, m.nb_nodes, 4, 1, 3
    # Yd = 0 weibull [1 MPa, 5]: the reference's engine and draw order (seed 2024)
    from akantu_amd import rng as akrng
    yd = akrng.weibull_thresholds(E, 1e6, 5.0, seed=2024)
    Eyoung, nu, rho, Sd, eps = 1e9, 0.3, 1200.0, 5e6, 0.05
    model = akx.SolidMechanicsModel(m)
    model.add_material(kind=akx.NEOHOOKEAN_DAMAGE, E=Eyoung, nu=nu, rho=rho, Sd=Sd, Yd=yd)
    model.init_full()
    X = m.nodes
    blocked = np.zeros_like(X, dtype=bool)
    blocked[np.isclose(X[:, 1], 0.0)] = True
    top = np.nonzero(np.isclose(X[:, 1], X[:, 1].max()))[0]
    model.set("blocked_dofs", blocked)
    dt = 0.8 * model.get_stable_time_step()
    model.set_time_step(dt)
    model.set_dirichlet_increment(top, 1, 20.0 * dt)
    v0 = np.zeros_like(X)
    reset(model, v0)
    log("E=%d N=%d dt=%.3e" % (E, N, dt))

    # training set and training on the GPU
    model.surrogate_setup(0, eps)
    t0 = time.perf_counter()
    model.surrogate_generate(40000, 250, inc_std=2e-4, yd_scale=1e6, yd_shape=5.0, seed=1)
    model.synchronize()
    t_gen = time.perf_counter() - t0
    iters = args.train_iters
    t0 = time.perf_counter()
    loss = model.surrogate_train(iters, batch=8192, lr=2e-3, seed=7)
    t_train = time.perf_counter() - t0
    log("training set 10M samples in %.2f s; %d Adam iterations (batch 8192) in %.2f s, "
        "final loss %.3e" % (t_gen, iters, t_train, loss))

    # plain vs surrogate step, each on the launch shape the autotuner picks
    # for it (the surrogate kernels write element- or node-ordered slots)
    res = {}
    for name, on in (("plain", False), ("surrogate", True)):
        model.surrogate_enable(on)
        model.set("damage", np.zeros(E))
        reset(model, v0)
        shape, _ = model.autotune(akx.MIXED, 5)
        model.set_launch(**shape)
        log("%s: autotuned launch %s" % (name, shape))
        model.set("damage", np.zeros(E))
        t = timed(model, akx.MIXED, args.steps, args.warmup, v0)
        el_ms, nd_ms = kernel_times(model, akx.MIXED, min(args.steps, 100), v0)
        res[name] = dict(ms_per_step=1e3 * t / args.steps, qp_per_s=E * Q * args.steps / t,
                         element_ms=el_ms, node_ms=nd_ms, launch=shape)
        if on:
            # the network kernel alone, HIP events around its launch in the
            # last step of the kernel-timing run
            res[name]["network_ms"] = model.query("surrogate_network_ms")
        log("%s: %.4f ms/step (element %.4f ms, node %.4f ms)" % (name, res[name]["ms_per_step"],
                                                                el_ms, nd_ms))
    # accuracy on the simulation's own strain states after `steps` plain steps
    steps = args.error_steps if args.error_steps != 1000 else 400
    model.surrogate_enable(False)
    model.set("damage", np.zeros(E))
    reset(model, v0)
    model.solve_step(steps, akx.MIXED)
    u_plain = model.get_displacement()
    G = model.get("grad_u").reshape(-1, 3, 3).transpose(0, 2, 1)  # (i, j) = du_i/dX_j
    dmg = model.get("damage")
    Eg = 0.5 * (G + G.transpose(0, 2, 1) + np.einsum("pki,pkj->pij", G, G))
    idx = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    Ev = np.stack([Eg[:, i, j] for i, j in idx], axis=1)
    lam, mu = nu * Eyoung / ((1 + nu) * (1 - 2 * nu)), Eyoung / (2 * (1 + nu))
    inside = np.all(np.abs(Ev) <= eps, axis=1)
    x = np.concatenate([Ev / eps, dmg[:, None]], axis=1).astype(np.float32)
    trE = Ev[:, 0] + Ev[:, 1] + Ev[:, 2]
    S_lin = 2 * mu * Ev
    S_lin[:, :3] += (lam * trE)[:, None]
    # plain law in FP64 from C = I + 2E
    C = np.zeros((E, 3, 3))
    for k, (i, j) in enumerate(idx):
        C[:, i, j] += 2 * Ev[:, k]
        if i != j:
            C[:, j, i] += 2 * Ev[:, k]
    C += np.eye(3)
    Ci = np.linalg.inv(C)
    lnJ = 0.5 * np.log(np.linalg.det(C))
    S_nh = mu * (np.eye(3) - Ci) + (lam * lnJ)[:, None, None] * Ci
    S_ref = np.stack([S_nh[:, i, j] for i, j in idx], axis=1)
    wgt = (1.0 - dmg)[:, None]
    strong = inside & (np.linalg.norm(S_ref, axis=1) > 1e-3 * np.linalg.norm(S_ref, axis=1).max())

    def stress_error():
        """The network's stress against the plain law on these strain
        states: relative over the points inside the envelope, and the 99th
        percentile per point."""
        y = model.surrogate_infer(x).astype(np.float64)
        diff = (S_lin + y[:, :6] * model.query("surrogate_stress_scale") - S_ref) * wgt
        rel = float(np.linalg.norm(diff[inside]) / np.linalg.norm((S_ref * wgt)[inside]))
        per_qp = np.linalg.norm(diff, axis=1) / np.maximum(np.linalg.norm(S_ref * wgt, axis=1),
                                                            1e-30)
        return rel, float(np.percentile(per_qp[strong], 99))

    def surrogate_run():
        """The surrogate run itself, from the same start: displacement error
        against the plain run and fallback count."""
        model.surrogate_enable(True)
        model.set("damage", np.zeros(E))
        reset(model, v0)
        model.surrogate_fallbacks(reset=True)
        model.solve_step(steps, akx.MIXED)
        u_sur = model.get_displacement()
        return (float(np.linalg.norm(u_sur - u_plain) / np.linalg.norm(u_plain)),
                model.surrogate_fallbacks())

    rel_stress, p99 = stress_error()
    rel_u, fb = surrogate_run()
    log("after %d steps: stress rel error %.3e (p99 per point %.3e), %.2f%% points inside the "
        "envelope, displacement rel error %.3e, fallbacks %d" %
        (steps, rel_stress, p99, 100 * inside.mean(), rel_u, fb))
    # the same network in BF16, the training format (FP16 above)
    model.surrogate_format("bf16")
    model.set_launch(**res["surrogate"]["launch"])
    model.set("damage", np.zeros(E))
    t = timed(model, akx.MIXED, args.steps, args.warmup, v0)
    el_ms, nd_ms = kernel_times(model, akx.MIXED, min(args.steps, 100), v0)
    bf = dict(ms_per_step=1e3 * t / args.steps, element_ms=el_ms, node_ms=nd_ms,
              network_ms=model.query("surrogate_network_ms"))
    bf["rel_stress_error"], bf["p99_point_rel_stress_error"] = stress_error()
    bf["rel_u_error_vs_plain"], _ = surrogate_run()
    model.surrogate_format("fp16")
    log("bf16 network: %.4f ms/step (network %.4f ms), stress rel error %.3e (p99 %.3e), "
        "displacement rel error %.3e" % (bf["ms_per_step"], bf["network_ms"],
                                         bf["rel_stress_error"], bf["p99_point_rel_stress_error"],
                                         bf["rel_u_error_vs_plain"]))
    flop_qp = 2 * (7 * 128 + 3 * 128 * 128 + 128 * 7)
    net_ms = res["surrogate"]["network_ms"]
    tflops = flop_qp * E * Q / (net_ms * 1e-3) / 1e12
    out = {
        "metric": "element-QP updates/s per explicit step, surrogate vs plain neo-Hookean+damage",
        "value": res["surrogate"]["qp_per_s"], "unit": "QP-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["surrogate"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 MFMA network (f32 accumulation; trained in bf16) / f32 element math / f64 nodal state",
        "data": "synthetic",
        "config": {"workload": "plate-surrogate: notched plate 0.1 x 0.1 x 0.02 m, %d "
                               "tetrahedron_4, %d nodes, neo-Hookean + Marigo damage E=1 GPa "
                               "nu=0.3 rho=1200 Sd=5 MPa, Yd ~ 1 MPa Weibull(m=5), tension at "
                               "20 m/s; 7-128x4-7 smoothstep-swish surrogate trained on 10M GPU-generated "
                               "samples" % (E, N),
                   "elements": E, "nodes": N, "qp_per_element": Q, "parallelism": "1 GPU"},
        "plain": res["plain"], "surrogate": res["surrogate"], "surrogate_bf16_network": bf,
        "surrogate_over_plain": res["surrogate"]["qp_per_s"] / res["plain"]["qp_per_s"],
        "training": {"samples": int(model.query("surrogate_samples")), "generate_s": t_gen,
                     "iterations": iters, "batch": 8192, "train_s": t_train, "final_loss": loss},
        "accuracy": {"steps": steps, "rel_stress_error": rel_stress,
                     "p99_point_rel_stress_error": p99, "inside_envelope": float(inside.mean()),
                     "rel_u_error_vs_plain": rel_u, "fallback_points": int(fb)},
        "roofline": {"bound": "mfma",
                     "kernel": "akx::(anonymous namespace)::k_sur_infer<akx::mlp::Fp16>(...)",
                     "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None,
                     "algorithmic_flops_per_qp": flop_qp, "avg_ms": net_ms},
        "cpu_baseline": None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cube-tet4",
                    choices=["cube-tet4", "hex8-neohookean", "ensemble", "plate-surrogate"])
    ap.add_argument("--train-iters", type=int, default=6000,
                    help="Adam iterations of the surrogate (plate-surrogate)")
    ap.add_argument("--n", type=int, default=100, help="hexes per cube edge")
    ap.add_argument("--error-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-strain-run", action="store_true",
                    help="hex8-neohookean: skip the run to 10 %% strain")
    ap.add_argument("--copy-only", action="store_true",
                    help="only the 1 GiB device copy (calibration of the traffic counters)")
    ap.add_argument("--launch-mixed", default=None,
                    help="element_block,node_block,xcd_remap,mixed_recompute,slot_layout,node_stage "
                         "(skips the autotuner for the mixed path)")
    ap.add_argument("--launch-fp64", default=None, help="same, for the FP64 path")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the launch shapes even where profiled ones are committed")
    args = ap.parse_args()
    if args.copy_only:
        print(json.dumps({"copy_bytes": 1 << 30, "copy_gbs": copy_bandwidth(1 << 30, 3)}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from akantu_amd import akx
    fn = {"cube-tet4": bench_cube, "hex8-neohookean": bench_hex8,
          "ensemble": bench_ensemble, "plate-surrogate": bench_plate_surrogate}[args.workload]
    print(json.dumps(fn(args, akx)), flush=True)


if __name__ == "__main__":
    main()
