"""bench.py's host-side inputs, on the CPU: the committed profile summaries it
reads (HBM bytes per launch, SQ instruction counts), the work model of the
roofline figure on hand counts, and the refusal of N > 1 GPUs."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_summary_is_read():
    t = bench.read_traffic()
    assert t is not None
    # the trajectories alone are 4 * 2 * 1000 * 65536 bytes; the blocked layout
    # writes each 32-byte sector once, so the launch moves little more
    assert 5.24e8 <= t <= 1.5 * 5.24288e8


def test_issue_fraction_from_counters():
    r = bench.read_issue(40.0)
    assert r is not None and 0.0 < r["frac"] < 1.0
    # the fraction scales inversely with the kernel time it is applied to
    assert bench.read_issue(20.0)["frac"] == pytest.approx(2.0 * r["frac"])


def test_missing_summaries_give_none(tmp_path):
    assert bench.read_traffic(str(tmp_path / "none.txt")) is None
    assert bench.read_issue(40.0, str(tmp_path / "none.txt")) is None


def test_work_model_on_hand_counts():
    """Lotka-Volterra: P = sum_j (1 + 2 order_j) = 3 + 5 + 3 = 11, N = 2,
    M = 3, nnz(V) = 4; each counter alone times its weight."""
    tabs = {"N": 2, "M": 3, "prop_order": np.array([[1, 1, 0], [0, 1, 1]]),
            "V": np.array([[0, 0, 1, 0], [0, 1, -1, 0], [1, 1, 1, 0], [1, 2, -1, 0]])}
    zero = dict(uniforms=0, ptrs_iterations=0, inversion_steps=0, asym_draws=0, exact_draws=0, leap_decisions=0,
                leaps=0, rejected_leaps=0, ssa_steps=0, sample_rows=0)
    per_unit = {"uniforms": 37, "ptrs_iterations": 62, "inversion_steps": 3, "asym_draws": 95, "exact_draws": 300,
                "leap_decisions": 11 + 6 + 92 + 5, "leaps": 8 + 2 + 1, "rejected_leaps": 8 + 2,
                "ssa_steps": 11 + 12 + 13 + 4 / 3, "sample_rows": 16}
    for k, w in per_unit.items():
        st = dict(zero)
        st[k] = 1000
        assert bench.work_model(tabs, 2, st) == pytest.approx(1000 * w), k


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
