"""Writes the checker's Taylor-Green series used by the GPU parity tests.

The C restatement (oracle/, pinned by tests/test_oracle.py) advances the
main-path Taylor-Green box (examples/Taylor_Green_Vortex deck: Re 1600,
Ma 0.1, 643c, alpha 0.49, rk3) at 64^3 with dt = 0.004 for 100 RK3 steps:

    python tests/golden/make_oracle_series.py

Output (data only): tgv64_main.npz with
  stats   (100, 3)  KE, enstrophy, dissipation at stage 1 of every step
  probe_* (9,9,9)   every 8th node of q1..q5, rho, u, v, w, prs, tmp after
                    the last step
  l2_*    scalar    root-mean-square of each final field over the interior
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

N, DT, STEPS = 64, 0.004, 100
NAMES = ["q1", "q2", "q3", "q4", "q5", "rho", "u", "v", "w", "prs", "tmp"]


def main():
    b = pyoracle.OracleBox(N, variant=0)
    b.init_tgv()
    stats = np.zeros((STEPS, 3))
    t0 = time.time()
    for s in range(STEPS):
        stats[s] = b.step(DT, stats=True)
        if s % 10 == 0:
            print(s, stats[s], f"{time.time() - t0:.1f}s", flush=True)
    out = {"stats": stats, "n": N, "dt": DT, "steps": STEPS}
    for nm in NAMES:
        f = b.get(nm)
        out["probe_" + nm] = f[::8, ::8, ::8].copy()
        out["l2_" + nm] = np.sqrt(np.mean(f * f))
    np.savez_compressed(os.path.join(HERE, "tgv64_main.npz"), **out)
    print("done", time.time() - t0)


if __name__ == "__main__":
    main()
