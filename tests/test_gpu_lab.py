"""The GPU error lab (tinysolverdx.lab): a quick sweep runs end to end and its
numbers obey the bounds the reference's own tests use."""
import csv

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_quick_lab_sweep(tmp_path):
    import tinysolverdx as tsdx
    from tinysolverdx import lab
    lab.main(["--out", str(tmp_path), "--quick", "--n", "8", "--n-tier", "32"])
    for f in ("error_vs_cond.csv", "error_vs_n.csv", "tiers_vs_cond.csv", "REPORT.md", "error_vs_cond_fp32.png",
              "backward_error_vs_cond.png", "backward_error_vs_n.png", "tiers_vs_cond.png"):
        assert (tmp_path / f).exists(), f
    rows = list(csv.DictReader(open(tmp_path / "error_vs_cond.csv")))
    for r in rows:
        u = tsdx.unit_roundoff(r["precision"])
        bw = float(r["backward_max"])
        if not np.isnan(bw):  # backward stable (test_solvers.cpp:58 uses 4 n u for Cholesky, 8 n u for LU)
            assert bw <= 8 * 8 * u, r
    tiers = list(csv.DictReader(open(tmp_path / "tiers_vs_cond.csv")))
    assert {r["tier"] for r in tiers} == set(lab.TIERS)
    for r in tiers:
        assert float(r["solve_backward_max"]) <= 1e-14, r
        assert int(r["info_gt0"]) == 0
    # fp32 refinement converges without the float64 solve while kappa * u_fp32 << 1
    easy = [r for r in tiers if r["tier"] == "fp32+IR" and float(r["kappa"]) <= 1e4]
    assert easy and all(float(r["fallback_frac"]) == 0.0 for r in easy)


def test_quick_bench_sweep(tmp_path):
    """tinysolverdx.bench: the per-size sweep writes the reference's CSV
    columns and the roofline chart, and its numbers sit under the measured
    ceilings."""
    from tinysolverdx import bench
    rows = bench.main(["--out", str(tmp_path), "--quick"])
    assert (tmp_path / "bench_fp32.csv").exists() and (tmp_path / "roofline_fp32.png").exists()
    got = list(csv.DictReader(open(tmp_path / "bench_fp32.csv")))
    assert {"kind", "n", "precision", "series", "gflops", "batch", "threads"} <= set(got[0])
    assert {r["series"] for r in got} == set(bench.SERIES)
    for r in rows:
        assert np.isfinite(r["gflops"]) and r["gflops"] > 0
        assert r["frac_of_measured_hbm"] < 1.5, r  # algorithmic bytes cannot beat the stream rate by much
