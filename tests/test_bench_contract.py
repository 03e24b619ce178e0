"""bench.py's contract on the CPU: one GPU only (--gpus above 1 exits with a
message before touching a device), and the byte models it reports rooflines
with follow the arrays each step touches."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_position_form_drops_velocity_and_acceleration():
    """The position form reads u and the previous displacement and writes u:
    3 N d 8 bytes less nodal state than u, v and a read and written."""
    E, N, n, Q, d = 1000, 300, 4, 1, 3
    _, with_va = bench.algorithmic_bytes(E, N, n, Q, d, 4)
    _, position = bench.algorithmic_bytes(E, N, n, Q, d, 4, position=True)
    assert with_va - position == 3 * N * d * 8


def test_fused_mixed_against_fused_fp64():
    """The mixed fused step differs from the FP64 one by the nodal state (the
    position form: the previous displacement read in place of v and a, and
    dt^2 / M in place of the mass and the per-node time step) and, on
    triangles, the
    32-byte record per evaluation in place of the element list, LDS slots,
    connectivity and coordinates (without records, the same streams as the
    FP64 step)."""
    E, N, d, ev, nblk = 10000, 5200, 2, 11700, 21
    fp64 = bench.fused_bytes(E, N, 3, d, ev, nblk)
    mixed = bench.fused_mixed_bytes(E, N, 3, d, ev, nblk, dt_per_node=True)
    # the FP64 step reads the mass; the position form dt^2 / M in place of
    # the mass and the per-node time step
    assert mixed - fp64 == pytest.approx(ev * 32 - ev * (4 + 6) - E * 3 * 4 - N * d * 8
                                         - 3 * N * d * 8)
    gathers = bench.fused_mixed_bytes(E, N, 3, d, ev, nblk, dt_per_node=True, records=False)
    assert gathers - fp64 == pytest.approx(-3 * N * d * 8)
    # tetrahedra keep the element lists and gather the coordinates
    fp64_t = bench.fused_bytes(E, N, 4, 3, ev, nblk)
    mixed_t = bench.fused_mixed_bytes(E, N, 4, 3, ev, nblk)
    assert fp64_t - mixed_t == pytest.approx(3 * N * 3 * 8)
