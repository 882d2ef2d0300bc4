"""bench.py's one JSON line, on a small Taylor-Green box and a small
Riemann box: the fields of the contract, a throughput consistent with the
step time, and a roofline object whose achieved rate is its bytes over the
measured launch time."""
import json
import os
import subprocess
import sys

import numpy as np
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
    d = run_bench("--n", "64", "--steps", "3", "--warmup", "1", "--no-cpu", "--legs", "plain")
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
    # the value is the median of the event-timed runs
    runs = d["timing"]["runs_ms"]
    assert len(runs) == 5
    assert abs(d["ms_per_step"] - float(np.median(runs)) / 3) <= 1e-9 * d["ms_per_step"]
    sr = d["step_roofline"]
    assert abs(sr["frac"] - d["value"] * 4224.0 / 8e12) <= 1e-12
    assert d["state"]["finite"]
    assert "batched" not in d and "adjoint" not in d


def test_line_carries_every_part_of_the_metric():
    """Plain, batched and adjoint parts in one line (small sizes)."""
    d = run_bench("--n", "32", "--steps", "2", "--warmup", "1", "--no-cpu", "--nbox", "3",
                  "--ens-sizes", "32,48", "--adj-n", "16", "--nsteps", "4", "--ckpt", "2",
                  "--fd-dirs", "2")
    assert KEYS <= set(d)
    b = d["batched"]
    assert len(b["boxes"]) == 3 and all(v in (32, 48) for bx in b["boxes"] for v in bx)
    assert abs(d["batched_over_loop"] - b["value"] / b["loop_value"]) <= 1e-12
    a = d["adjoint"]
    assert abs(d["adjoint_over_forward"] -
               (a["recorded_forward_ms"] + a["reversal_ms"]) / a["forward_ms"]) <= 1e-9
    assert a["fd_directions"] == 2 and a["fd_max_rel_err"] < 1e-5


def test_riemann_line():
    d = run_bench("--workload", "riemann2d", "--n", "256", "--steps", "3", "--warmup", "1",
                  "--no-cpu")
    assert KEYS <= set(d)
    assert abs(d["value"] - 257 ** 2 / (d["ms_per_step"] * 1e-3)) <= 1e-6 * d["value"]
    # the flop count comes from the PMC summary of the 2048^2 workload only
    assert d["roofline"] is None
    assert d["cpu_baseline"] is None
