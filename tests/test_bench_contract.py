"""The measurement logic of bench.py, checked without a GPU: the algorithmic
bytes behind the roofline figures, the traffic figures read from the
committed profile, the roofline arithmetic and the one-GPU contract."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_fused_update_bytes_per_cell():
    # per stage: state in and out (80 B), x and y fluxes (2 x 5 words);
    # stages 2 and 3 also read the base state (2/3 x 40 B on average)
    assert bench.update_bytes_per_cell("split") == pytest.approx(80 + 40 + 80 / 3)
    assert bench.update_bytes_per_cell("fp64") == pytest.approx(80 + 80 + 80 / 3)
    assert round(bench.update_bytes_per_cell("split"), 1) == 146.7


def test_profiled_traffic_comes_from_the_committed_profile():
    split = bench.profiled_traffic("k_zupdate", "split")
    fp64 = bench.profiled_traffic("k_zupdate", "fp64")
    cells = 256 ** 3
    # measured HBM bytes per launch lie near the algorithmic bytes (re-reads
    # of halo cells and tile edges make them a little larger)
    for t, prec in ((split, "split"), (fp64, "fp64")):
        algo = bench.update_bytes_per_cell(prec) * cells
        assert t is not None and 0.95 * algo < t < 1.25 * algo
    assert bench.profiled_traffic("no_such_kernel", "split") is None


def test_step_roofline_arithmetic():
    r = bench.step_roofline(4.0e9)
    assert r["achieved"] == pytest.approx(4.0e9 * 320 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s"


def test_published_peaks():
    assert bench.HBM_PEAK_GBS == 8000.0
    assert bench.VALU_FP32_TFLOPS == 157.3 and bench.VALU_FP64_TFLOPS == 78.6


def test_host_cpu_names_the_host():
    h = bench.host_cpu()
    assert h["nproc"] >= 1 and isinstance(h["model"], str)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
