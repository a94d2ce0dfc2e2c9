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



def test_bench_precision_form_line():
    """bench.py --form precision: one JSON line with the error rows at each mark
    for FP32 storage with and without the composition offset, and the rates."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--form", "precision",
                        "--size", "256", "--precision-marks", "10", "30"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert d["form"] == "precision" and d["grid"] == [256, 256]
    assert [row["step"] for row in d["rows"]] == [10, 30]
    for row in d["rows"]:
        for key in ("max_abs_dx_Cr", "max_abs_dx_Nb", "max_abs_dx_Cr_offset",
                    "max_abs_dx_Nb_offset", "mass_drift_Cr_32", "mass_drift_Cr_32_offset"):
            assert key in row, key
        assert 0 < row["max_abs_dx_Cr_offset"] < row["max_abs_dx_Cr"] < 1e-5
    rates = d["rates"]
    for key in ("fp64", "fp32_storage", "fp32_storage_offset"):
        assert rates[key] > 0, key
    assert rates["steps"] == 30 and len(rates["offset"]) == 2
