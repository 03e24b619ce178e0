"""Independent specimens advancing together in one model (per-node time
steps, per-material stable time steps) and the graph-captured step loop.

Each specimen of the ensemble is its own SolidMechanicsModel in the reference
(examples/explicit/explicit_dynamic.cc:38-106 run once per specimen).  No
element joins two specimens, so every node's force gather, lumped mass and
update only see its own specimen, in the same order as in a model of the
specimen alone: the FP64 ensemble must equal, bit for bit, the oracle run on
each specimen separately with that specimen's own stable time step.
"""
import numpy as np
import pytest

from akantu_amd import akx
from akantu_amd import mesh as akm
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _ensemble(count=6, seed=7):
    sp = akm.ensemble_tri3(count, min_elems=200, max_elems=900, seed=seed)
    return akm.concatenate_specimens(sp)


def _gpu_model(en):
    model = akx.SolidMechanicsModel(en.mesh)
    for s in en.specimens:
        model.add_material(E=s.E, nu=s.nu, rho=s.rho)
    model.assign_materials(en.element_material)
    model.init_full()
    model.set("blocked_dofs", en.blocked)
    model.set("velocity", en.velocity)
    dt_mat = model.get_stable_time_step_per_material()
    model.set_node_time_steps(0.8 * dt_mat[en.node_specimen()])
    return model, dt_mat


def _oracle_specimen(en, k, dt, steps):
    s = en.specimens[k]
    m = en.specimen_mesh(k)
    cpu = orc.OracleModel(m, [dict(E=s.E, nu=s.nu, rho=s.rho)])
    cpu.blocked[:] = s.blocked
    cpu.nodal("v")[:] = s.velocity
    stable = cpu.stable_time_step()
    cpu.set_time_step(dt)
    cpu.solve_step(steps)
    return stable, cpu.nodal("u").copy()


@pytest.mark.parametrize("graph,layout", [(False, 0), (True, 0), (False, 6), (True, 6)])
def test_ensemble_fp64_equals_each_specimen_alone(graph, layout):
    """One specimen per material, one time step per node; layout 6 is the
    fused node-block step."""
    en = _ensemble()
    model, dt_mat = _gpu_model(en)
    model.set_launch(slot_layout=layout)
    steps = 200
    if graph:
        model.solve_step_graph(steps, akx.FP64)
    else:
        model.solve_step(steps, akx.FP64)
    u = model.get_displacement()
    for k in range(len(en.specimens)):
        stable, ref = _oracle_specimen(en, k, 0.8 * dt_mat[k], steps)
        assert dt_mat[k] == stable, k
        n0, n1 = en.node_offsets[k], en.node_offsets[k + 1]
        assert np.array_equal(u[n0:n1], ref), k


@pytest.mark.parametrize("layout", [0, 6])
def test_ensemble_mixed_error_per_specimen(layout):
    en = _ensemble(count=5, seed=11)
    model, dt_mat = _gpu_model(en)
    model.set_launch(slot_layout=layout)
    steps = 1000
    model.solve_step_graph(steps, akx.MIXED)
    u = model.get_displacement()
    for k in range(len(en.specimens)):
        _, ref = _oracle_specimen(en, k, 0.8 * dt_mat[k], steps)
        n0, n1 = en.node_offsets[k], en.node_offsets[k + 1]
        err = np.linalg.norm(u[n0:n1] - ref) / np.linalg.norm(ref)
        assert err < 1e-5, (k, err)


def test_fused_mixed_records_equal_gathers():
    """The mixed fused step on triangles rebuilds B from one 32-byte record
    per element evaluation (mixed_recompute 1: node ids, LDS slots and the
    material packed together) or from coordinates gathered through element
    ids (0): the same displacements bit for bit, on 300 specimens (material
    indices past 8 bits, in the record's second and third words), and each
    specimen within 1e-5 of the oracle."""
    en = _ensemble(count=300, seed=5)
    out = []
    for rc in (1, 0):
        model, dt_mat = _gpu_model(en)
        assert model.query("fused_available") == 1 and model.query("fused_records") == 1
        model.set_launch(mixed_recompute=rc, slot_layout=6)
        model.solve_step(300, akx.MIXED)
        out.append(model.get_displacement())
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0]).max() > 0
    for k in (0, 255, 256, 299):
        _, ref = _oracle_specimen(en, k, 0.8 * dt_mat[k], 300)
        n0, n1 = en.node_offsets[k], en.node_offsets[k + 1]
        assert np.linalg.norm(out[0][n0:n1] - ref) / np.linalg.norm(ref) < 1e-5, k


@pytest.mark.parametrize("precision,layout", [(akx.FP64, 0), (akx.FP64, 6), (akx.MIXED, 0),
                                              (akx.MIXED, 4), (akx.MIXED, 5), (akx.MIXED, 6)])
def test_graph_replays_match_eager(precision, layout):
    """Replays of one captured graph, with host work and synchronisations in
    between, give what the eager launches give, every time; changing the time
    step re-captures.  Layouts 4 and 5: the patch kernel and the per-patch
    partial sums inside the graph."""
    m = akm.bar_tri3(60, 12)
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 0], 0.0)] = True
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.01 * np.sin(np.pi * X[:, 0] / 2)
    models = []
    for _ in range(2):
        g = akx.SolidMechanicsModel(m)
        g.add_material(E=210e9, nu=0.3, rho=7800.0)
        g.init_full()
        g.set("blocked_dofs", mask)
        g.set("velocity", v0)
        g.set_time_step(0.8 * g.get_stable_time_step())
        g.set_launch(slot_layout=layout)
        models.append(g)
    eager, graph = models
    for r in range(4):
        if r == 2:
            dt = 0.5 * eager.get_stable_time_step()
            eager.set_time_step(dt)
            graph.set_time_step(dt)
        eager.solve_step(25, precision)
        graph.solve_step_graph(25, precision)
        ue, ug = eager.get_displacement(), graph.get_displacement()
        assert np.array_equal(ue, ug), r
        assert np.array_equal(eager.get_velocity(), graph.get_velocity()), r
        # host work between replays
        assert eager.get_energy("kinetic") == graph.get_energy("kinetic")


@pytest.mark.parametrize("precision,layout", [(akx.FP64, 0), (akx.FP64, 6), (akx.MIXED, 5),
                                              (akx.MIXED, 6)])
def test_graph_captured_again_after_reinit(precision, layout):
    """akx_init frees and rebuilds every derived device buffer (CSR, patches,
    node blocks, element records): a graph captured before it is dropped and
    the next graph step captures again, equal to eager launches bit for bit."""
    m = akm.bar_tri3(60, 12)
    X = m.nodes
    mask = np.zeros_like(X, dtype=bool)
    mask[np.isclose(X[:, 0], 0.0)] = True
    v0 = np.zeros_like(X)
    v0[:, 0] = 0.01 * np.sin(np.pi * X[:, 0] / 2)
    models = []
    dt = None
    for _ in range(2):
        g = akx.SolidMechanicsModel(m)
        g.add_material(E=210e9, nu=0.3, rho=7800.0)
        g.init_full()
        g.set("blocked_dofs", mask)
        g.set("velocity", v0)
        dt = 0.8 * g.get_stable_time_step()
        g.set_time_step(dt)
        g.set_launch(slot_layout=layout)
        models.append(g)
    eager, graph = models
    eager.solve_step(25, precision)
    graph.solve_step_graph(25, precision)
    for g in models:
        g.init_full()
        # the very same time step and launch shape: every key of the graph
        # captured before is unchanged, so only the re-initialisation itself
        # can keep the stale graph (whose buffers are freed) from replaying
        g.set_time_step(dt)
        g.set_launch(slot_layout=layout)
    eager.solve_step(25, precision)
    graph.solve_step_graph(25, precision)
    assert np.array_equal(eager.get_displacement(), graph.get_displacement())
    assert np.array_equal(eager.get_velocity(), graph.get_velocity())
