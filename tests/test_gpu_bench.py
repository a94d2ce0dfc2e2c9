"""bench.py keeps its output contract: one JSON line with the fields the
benchmark driver reads (a small plain run, no CPU leg)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--plain-only", "--no-cpu",
                        "--size", "512", "--steps", "30", "--warmup", "6"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 30 and d["warmup"] == 6
    assert d["value"] > 0 and d["higher_is_better"] is True and d["dtype"] == "f64"
    assert d["value"] == pytest.approx(512 * 512 / (d["ms_per_step"] / 1e3), rel=1e-9)
    assert "workload" in d["config"]
    rf = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf, key
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-12)
    assert d["cpu_baseline"] is None
    # the timed loop and the kernel leg ran in the same clock state
    t = d["timing"]
    assert t["four_step_kernel_us_per_step"] <= 1.10 * t["timed_loop_us_per_step"]
    assert t["kernel_leg_over_timed_loop"] == pytest.approx(
        t["four_step_kernel_us_per_step"] / t["timed_loop_us_per_step"], rel=1e-9)
    # the dominant kernel runs four steps; the shorter ones are reported beside it
    assert rf["cell_updates_per_launch"] == 4 * 512 * 512
    for key in ("three_step_kernel", "two_step_kernel", "one_step_kernel"):
        assert rf[key]["kernel_avg_us"] > 0, key



def test_bench_precision_form_line():
    """bench.py --form precision: one JSON line with, for the no-precipitate
    and the seeded-precipitate workloads, the error rows at each mark (FP32
    storage with and without the composition offset against FP64) and the
    rates."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--form", "precision",
                        "--precision-size", "256", "--precision-marks", "10", "30",
                        "--precision-particles", "2"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])["precision"]
    assert d["grid"] == [256, 256]
    plain, parts = d["no_precipitates"], d["with_precipitates"]
    assert [row["step"] for row in plain["rows"]] == [10, 30]
    for row in plain["rows"]:
        for key in ("max_abs_dx_Cr_fp32", "max_abs_dx_Nb_fp32", "max_abs_dx_Cr_fp32_offset",
                    "max_abs_dx_Nb_fp32_offset", "mass_drift_Cr_fp32",
                    "mass_drift_Cr_fp32_offset", "mass_drift_Cr_fp64"):
            assert key in row, key
        assert 0 < row["max_abs_dx_Cr_fp32_offset"] < row["max_abs_dx_Cr_fp32"] < 1e-5
    for key in ("fp64", "fp32", "fp32_offset"):
        assert plain["rates_cell_updates_per_s"][key] > 0, key
    assert plain["steps"] == 30 and len(plain["offsets"]["fp32_offset"]) == 2
    # with precipitates the fractions are not zero and the FP32 error is small
    for row in parts["rows"]:
        assert row["delta_frac_fp64"] > 1e-3 and row["laves_frac_fp64"] > 1e-3
        assert row["phase_frac_abs_err_fp32_offset"] < 1e-6
