"""The benchmark's output contract and its operation accounting (CPU only)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import aderdg  # noqa: E402
import bench  # noqa: E402


def test_roofline_object_keys_and_arithmetic():
    r = bench.roofline_of(154576.0, 1000, 2.0, "k_predictor_lean<BnPDE,3>", 1.0e9,
                          bench.skipped_for("k_predictor_lean<BnPDE,3>", "bn", 3))
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in r
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and r["peak"] == bench.FP64_PEAK_TFLOPS
    assert r["achieved"] == pytest.approx(154576.0 * 1000 / 2.0e-3 / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    assert 0.0 < r["executed_achieved"] < r["achieved"]


@pytest.mark.parametrize("pde", ["euler", "bn"])
def test_first_sweep_skipped_is_a_part_of_the_count(pde):
    """The operations the time-line predictors skip are fewer than N of the N+1 sweeps'
    worth of the whole count, and grow with N."""
    prev = 0.0
    for N in (1, 2, 3, 4):
        total, _ = aderdg.flops_per_cell(N, pde)
        sk = bench.first_sweep_skipped(pde, N)
        assert 0.0 < sk < total * N / (N + 1)
        assert sk > prev
        prev = sk


def test_skipped_only_for_time_line_kernels():
    assert bench.skipped_for("k_predictor_nodes<.., 64><BnPDE,1>", "bn", 1) == 0.0
    assert bench.skipped_for("k_predictor<EulerPDE,4>", "euler", 4) == bench.first_sweep_skipped("euler", 4)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
