"""Writes the checker's Taylor-Green series used by the GPU parity tests.

The C restatement (oracle/, pinned by tests/test_oracle.py) advances the
main-path Taylor-Green box (examples/Taylor_Green_Vortex deck: Re 1600,
Ma 0.1, 643c, alpha 0.49, rk3) for 100 RK3 steps with dt = 0.004 * 64 / N,
the same acoustic CFL number at every size, (|u| + 1/Ma) dt / dx = 0.45:

    python tests/golden/make_oracle_series.py          # 64^3,  dt 0.004 -> tgv64_main.npz
    python tests/golden/make_oracle_series.py 256      # 256^3, dt 0.001 -> tgv256_main.npz

dt = 0.004 at 256^3 would be an acoustic CFL number of 1.8, past the RK3
stability limit of the compact scheme: the run grows and turns to NaN
within 20 steps.

Output (data only), tgv<N>_main.npz with
  stats   (100, 3)  KE, enstrophy, dissipation at stage 1 of every step
  probe_* (9,9,9)   every (N/8)-th node of q1..q5, rho, u, v, w, prs, tmp
                    after the last step
  l2_*    scalar    root-mean-square of each final field over the interior
and, for N > 64, three whole lines of every final field through the node
(N/8*3+1, N/8*5+2, N/8*7+3):
  xline_* (N+1,)    all i at that (j, k)
  yline_* (N+1,)    all j at that (i, k)
  zline_* (N+1,)    all k at that (i, j)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

STEPS = 100


def series_dt(n):
    """dt of the n^3 series: 0.004 at 64^3, scaled with the spacing."""
    return 0.004 * 64 / n
NAMES = ["q1", "q2", "q3", "q4", "q5", "rho", "u", "v", "w", "prs", "tmp"]


def line_node(n):
    """The node the whole-line probes pass through (i, j, k)."""
    s = n // 8
    return 3 * s + 1, 5 * s + 2, 7 * s + 3


def main(n=64):
    dt = series_dt(n)
    b = pyoracle.OracleBox(n, variant=0)
    b.init_tgv()
    stats = np.zeros((STEPS, 3))
    t0 = time.time()
    for s in range(STEPS):
        stats[s] = b.step(dt, stats=True)
        if s % 10 == 0:
            print(s, stats[s], f"{time.time() - t0:.1f}s", flush=True)
    out = {"stats": stats, "n": n, "dt": dt, "steps": STEPS}
    st = n // 8
    i0, j0, k0 = line_node(n)
    for nm in NAMES:
        f = b.get(nm)  # f[i, j, k] over nodes 0..n
        out["probe_" + nm] = f[::st, ::st, ::st].copy()
        out["l2_" + nm] = np.sqrt(np.mean(f * f))
        if n > 64:
            out["xline_" + nm] = f[:, j0, k0].copy()
            out["yline_" + nm] = f[i0, :, k0].copy()
            out["zline_" + nm] = f[i0, j0, :].copy()
    np.savez_compressed(os.path.join(HERE, f"tgv{n}_main.npz"), **out)
    print("done", time.time() - t0)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 64)
