"""bench.py's one JSON line, on a small Taylor-Green box and a small
Riemann box: the fields of the contract, a throughput consistent with the
step time, and a roofline object whose achieved rate is its bytes over the
measured launch time."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
        "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config", "roofline",
        "cpu_baseline"}


def run_bench(*args):
    print(f"case bench {' '.join(args)}", flush=True)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1, out.stdout
    return json.loads(lines[0])


def test_taylor_green_line():
    d = run_bench("--n", "64", "--steps", "3", "--warmup", "1", "--no-cpu")
    assert KEYS <= set(d)
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["dtype"] == "f64" and d["higher_is_better"] is True
    assert "workload" in d["config"]
    # grid-point-RK3-steps/s = n^3 / seconds per step
    assert abs(d["value"] - 64 ** 3 / (d["ms_per_step"] * 1e-3)) <= 1e-6 * d["value"]
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) <= 1e-12
    assert abs(r["achieved"] - r["algo_bytes_per_launch"] / (r["avg_launch_ms"] * 1e-3) / 1e9) \
        <= 1e-9 * r["achieved"]
    assert 0.0 < r["frac"] < 1.0
    assert d["cpu_baseline"] is None


def test_riemann_line():
    d = run_bench("--workload", "riemann2d", "--n", "256", "--steps", "3", "--warmup", "1",
                  "--no-cpu")
    assert KEYS <= set(d)
    assert abs(d["value"] - 257 ** 2 / (d["ms_per_step"] * 1e-3)) <= 1e-6 * d["value"]
    # the flop count comes from the PMC summary of the 2048^2 workload only
    assert d["roofline"] is None
    assert d["cpu_baseline"] is None
