"""Writes the golden line vectors of the reference's own modules.

Run in the container that holds the reference tree, after `make -C oracle ref`
has compiled src/{constdef,commtype,commvar,commfunc,singleton,derivative,
filter}.F90 into oracle/_ref/librefops.so:

    python tests/golden/make_golden.py

Outputs (data only, committed):
  lines_ref.npz     inputs and outputs of df_compact, diff6ec and
                    compact_filter (src/derivative.F90, src/filter.F90) for
                    ntype 1..4 and dim 64/128/256, the pre-factored LHS
                    tables and the filter coefficient tables
  state.ref_128.txt the reference-held series miniapps/tgv_solver_3d/
                    state.ref_128 (nstep, time, KE, enstrophy of the
                    128^3 mini-app), copied as a data fixture
"""
import ctypes
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

REF = "/root/reference"


def main():
    ref = pyoracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/librefops.so missing: run `make -C oracle ref`")
    rng = np.random.default_rng(20251202)
    out = {}
    for dim in (64, 128, 256):
        x = np.arange(-5, dim + 6) * (2 * np.pi / dim)
        for nt in (1, 2, 3, 4):
            f = rng.standard_normal(dim + 11)
            g = np.sin(3 * x) + 0.25 * np.cos(7 * x + 0.3)
            for tag, inp in (("rand", f), ("smooth", g)):
                k = f"{tag}_{nt}_{dim}"
                out["f_" + k] = inp
                out["df_" + k] = pyoracle.df_compact(nt, inp, lib=ref)
                out["d6_" + k] = pyoracle.diff6ec(nt, inp, lib=ref)
                out["ff098_" + k] = pyoracle.compact_filter(nt, inp, 0.49, 1.11, 0.98, lib=ref)
                out["ff109_" + k] = pyoracle.compact_filter(nt, inp, 0.49, 1.11, 1.09, lib=ref)
            n = ctypes.c_int()
            ac = np.zeros(3 * (dim + 7))
            ref.ref_fd_ac(nt, dim, ctypes.byref(n), ac.ctypes.data_as(pyoracle._D))
            out[f"fdac_{nt}_{dim}"] = ac[:3 * n.value].reshape(n.value, 3).copy()
            ref.ref_filter_ac(nt, dim, 0.49, ctypes.byref(n), ac.ctypes.data_as(pyoracle._D))
            out[f"flac_{nt}_{dim}"] = ac[:3 * n.value].reshape(n.value, 3).copy()
    out["coef_098"] = pyoracle.filter_coef(0.49, 1.11, 0.98, lib=ref)
    out["coef_109"] = pyoracle.filter_coef(0.49, 1.11, 1.09, lib=ref)
    # Thomas on small systems with random coefficients (|a|,|c| < 0.45)
    for n in (3, 5, 8, 16):
        a = rng.uniform(-0.45, 0.45, n)
        c = rng.uniform(-0.45, 0.45, n)
        d = rng.standard_normal(n)
        out[f"thomas_a_{n}"], out[f"thomas_c_{n}"], out[f"thomas_d_{n}"] = a, c, d
        out[f"thomas_x_{n}"] = pyoracle.thomas(a, c, d, lib=ref)
    np.savez_compressed(os.path.join(HERE, "lines_ref.npz"), **out)
    shutil.copyfile(os.path.join(REF, "miniapps", "tgv_solver_3d", "state.ref_128"),
                    os.path.join(HERE, "state.ref_128.txt"))
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
