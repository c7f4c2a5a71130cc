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


def test_algorithmic_bytes_model():
    """Per cell-step byte counts of the kernel table: Baer-Nunziato N=3 update reads uh and
    lduh (2 x 144 doubles), four face lines (4 x 36) and the ranges (18), writes uh' (144),
    the next ranges (18) and the time-step candidate (1)."""
    b = bench.algorithmic_bytes("bn", 3)
    assert b["update"] == 8 * (2 * 144 + 4 * 36 + 18) + 8 * (144 + 18 + 1)
    assert b["face"] == 2 * 8 * (2 * 2 * 36) + 2 * 8 * (2 * 36)
    e = bench.algorithmic_bytes("euler", 4)
    assert e["face"] == 2 * 8 * (2 * 2 * 20) + 2 * 8 * 20


def test_kernel_table_and_measured_peaks(monkeypatch):
    monkeypatch.setattr(bench, "PEAKS", (56.6, 6300.0, 4800.0))
    t = bench.kernel_table("bn", {3: dict(cells=1000, update=2.0, face=1.0)}, bench.PEAKS)
    u = t["3"]["update"]
    assert u["algorithmic_gbs"] == pytest.approx(bench.algorithmic_bytes("bn", 3)["update"] * 1000 / 2e-3 / 1e9)
    assert u["frac_hbm_peak"] == pytest.approx(u["algorithmic_gbs"] / bench.HBM_PEAK_GBS)
    assert u["frac_measured_read_peak"] == pytest.approx(u["algorithmic_gbs"] / 6300.0)
    r = bench.roofline_of(154576.0, 1000, 2.0, "k_predictor_lean<BnPDE,3>")
    assert r["frac_measured"] == pytest.approx(r["achieved"] / 56.6) and r["bound"] == "mfma"


def test_pinned_variant_spec():
    """ADERDG_PREDICTOR_VARIANT: one variant for every degree, or a per-degree list."""
    assert bench.pinned_variant("", 3) is None
    assert bench.pinned_variant("4", 2) == 4
    assert bench.pinned_variant("2:4,3:5", 3) == 5
    assert bench.pinned_variant("2:4,3:5", 2) == 4
    assert bench.pinned_variant("2:4,3:5", 4) is None
