"""ctypes binding of the checker libraries.

TEST INFRASTRUCTURE ONLY: importable by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product never imports it.

  liboracle.so        the C restatement (oracle.c)
  _ref/librefops.so   the reference's own line-operator modules compiled
                      from source with a bind(c) harness (built only where
                      the reference tree exists)
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HM = 5
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
_V = ctypes.c_void_p
_F = ctypes.c_double


def _p(a):
    return a.ctypes.data_as(_D)


_orc = None
_ref = None


def oracle_lib():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.orc_box_new.restype = _V
        lib.orc_box_new.argtypes = [_I, _I, _I, _I] + [_F] * 9
        lib.orc_box_free.argtypes = [_V]
        lib.orc_box_init_tgv.argtypes = [_V]
        lib.orc_box_step.argtypes = [_V, _F, _D]
        for n in ("orc_box_filterq", "orc_box_qswap", "orc_box_rhscal", "orc_box_updatefvar"):
            getattr(lib, n).argtypes = [_V]
        lib.orc_box_stats.argtypes = [_V, _D]
        lib.orc_box_get.argtypes = [_V, _I, _D]
        lib.orc_box_set.argtypes = [_V, _I, _D]
        lib.orc_box_get_deriv.argtypes = [_V, _I, _D]
        lib.orc_box_get_haloed.argtypes = [_V, _I, _D]
        lib.orc_box_set_sutherland.argtypes = [_V, _F]
        lib.orc_df_compact.argtypes = [_I, _I, _D, _D]
        lib.orc_diff6ec.argtypes = [_I, _I, _D, _D]
        lib.orc_compact_filter.argtypes = [_I, _I, _F, _F, _F, _D, _D]
        lib.orc_filter_coef.argtypes = [_F, _F, _F, _D]
        lib.orc_thomas_prep.argtypes = [_D, _D, _I, _D]
        lib.orc_thomas_solve.argtypes = [_D, _D, _I, _D]
        lib.orc_fd_init.argtypes = [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _D, _D, _D]
        lib.orc_filter_init.argtypes = [_I, _I, _F, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                        _D, _D, _D]
        lib.orc_omp_threads.restype = _I
        _orc = lib
    return _orc


def ref_lib():
    """The reference's compiled modules, or None where they were not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "librefops.so")
        if not os.path.exists(path):
            return None
        lib = ctypes.CDLL(path)
        lib.ref_df_compact.argtypes = [_I, _I, _D, _D]
        lib.ref_diff6ec.argtypes = [_I, _I, _D, _D]
        lib.ref_compact_filter.argtypes = [_I, _I, _F, _F, _F, _D, _D]
        lib.ref_filter_coef.argtypes = [_F, _F, _F, _D]
        lib.ref_thomas.argtypes = [_I, _D, _D, _D, _D]
        lib.ref_fd_ac.argtypes = [_I, _I, ctypes.POINTER(_I), _D]
        lib.ref_filter_ac.argtypes = [_I, _I, _F, ctypes.POINTER(_I), _D]
        _ref = lib
    return _ref


# ------------------------------------------------------------ line operators

def df_compact(ntype: int, f: np.ndarray, lib=None) -> np.ndarray:
    """f has dim+1+2*HM values (nodes -HM..dim+HM); returns df(0:dim)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    dim = f.size - 1 - 2 * HM
    df = np.zeros(dim + 1)
    fn = "orc_df_compact" if lib is None else "ref_df_compact"
    getattr(lib or oracle_lib(), fn)(ntype, dim, _p(f), _p(df))
    return df


def diff6ec(ntype: int, f: np.ndarray, lib=None) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    dim = f.size - 1 - 2 * HM
    df = np.zeros(dim + 1)
    fn = "orc_diff6ec" if lib is None else "ref_diff6ec"
    getattr(lib or oracle_lib(), fn)(ntype, dim, _p(f), _p(df))
    return df


def compact_filter(ntype: int, f: np.ndarray, alfa=0.49, beta_halo=1.11, beta_bound=0.98,
                   lib=None) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    dim = f.size - 1 - 2 * HM
    ff = np.zeros(dim + 1)
    fn = "orc_compact_filter" if lib is None else "ref_compact_filter"
    getattr(lib or oracle_lib(), fn)(ntype, dim, alfa, beta_halo, beta_bound, _p(f), _p(ff))
    return ff


def filter_coef(alfa=0.49, beta_halo=1.11, beta_bound=0.98, lib=None) -> np.ndarray:
    c = np.zeros(115)
    fn = "orc_filter_coef" if lib is None else "ref_filter_coef"
    getattr(lib or oracle_lib(), fn)(alfa, beta_halo, beta_bound, _p(c))
    return c


def thomas(a, c, d, lib=None) -> np.ndarray:
    """Pre-factor (a, c) and solve for right-hand side d (unit diagonal)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    d = np.array(d, dtype=np.float64)
    n = a.size
    x = np.zeros(n)
    if lib is not None:
        lib.ref_thomas(n, _p(a), _p(c), _p(d), _p(x))
        return x
    o = oracle_lib()
    ac = np.zeros(3 * n)
    o.orc_thomas_prep(_p(a), _p(c), n, _p(ac))
    o.orc_thomas_solve(_p(ac), _p(d), n, _p(x))
    return x


def line_system(op: str, ntype: int, dim: int, alfa=0.49):
    """(first, last, a, c, ac) of the oracle's derivative ('deriv') or filter
    ('filter') line system."""
    o = oracle_lib()
    nmax = dim + 7
    a = np.zeros(nmax); c = np.zeros(nmax); ac = np.zeros(3 * nmax)
    fi = _I(); la = _I()
    if op == "deriv":
        n = o.orc_fd_init(ntype, dim, ctypes.byref(fi), ctypes.byref(la), _p(a), _p(c), _p(ac))
    else:
        n = o.orc_filter_init(ntype, dim, alfa, ctypes.byref(fi), ctypes.byref(la),
                              _p(a), _p(c), _p(ac))
    return fi.value, la.value, a[:n].copy(), c[:n].copy(), ac[:3 * n].reshape(n, 3).copy()


# ------------------------------------------------------------ periodic box

FIELDS = {"q1": 0, "q2": 1, "q3": 2, "q4": 3, "q5": 4, "rho": 5, "u": 6, "v": 7, "w": 8,
          "prs": 9, "tmp": 10}


class OracleBox:
    """Periodic box of the C restatement; variant 0 = main path, 1 = mini-app."""

    def __init__(self, n, variant=0, reynolds=1600.0, mach=0.1, prandtl=0.72, gamma=1.4,
                 ref_tem=273.15, alfa=0.49, lengths=(2 * math.pi,) * 3):
        if isinstance(n, int):
            n = (n, n, n)
        self.n = tuple(int(v) for v in n)
        self.lib = oracle_lib()
        self.h = self.lib.orc_box_new(self.n[0], self.n[1], self.n[2], variant, reynolds, mach,
                                      prandtl, gamma, ref_tem, alfa, *lengths)

    def close(self):
        if self.h:
            self.lib.orc_box_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.n[0] + 1, self.n[1] + 1, self.n[2] + 1)

    def init_tgv(self):
        self.lib.orc_box_init_tgv(self.h)

    def get(self, name) -> np.ndarray:
        out = np.zeros(self.shape, order="F")
        self.lib.orc_box_get(self.h, FIELDS[name], _p(out))
        return out

    def get_haloed(self, name) -> np.ndarray:
        s = tuple(v + 2 * HM for v in self.shape)
        out = np.zeros(s, order="F")
        self.lib.orc_box_get_haloed(self.h, FIELDS[name], _p(out))
        return out

    def set(self, name, a):
        a = np.asfortranarray(a, dtype=np.float64)
        self.lib.orc_box_set(self.h, FIELDS[name], _p(a))

    def deriv(self, which) -> np.ndarray:
        out = np.zeros(self.shape, order="F")
        self.lib.orc_box_get_deriv(self.h, which, _p(out))
        return out

    def step(self, dt, stats=False):
        st = np.zeros(3)
        self.lib.orc_box_step(self.h, dt, _p(st) if stats else None)
        return tuple(st) if stats else None

    def filterq(self):
        self.lib.orc_box_filterq(self.h)

    def qswap(self):
        self.lib.orc_box_qswap(self.h)

    def rhscal(self):
        self.lib.orc_box_rhscal(self.h)

    def stats(self):
        st = np.zeros(3)
        self.lib.orc_box_stats(self.h, _p(st))
        return tuple(st)

    def state_haloed(self):
        """Interior state as haloed host arrays (halos zero) in the layout of
        astr_gpu_amd's upload_state."""
        nx, ny, nz = self.shape
        s = (nx + 2 * HM, ny + 2 * HM, nz + 2 * HM)
        I = (slice(HM, HM + nx), slice(HM, HM + ny), slice(HM, HM + nz))
        out = {"q": np.zeros(s + (5,), order="F"), "rho": np.zeros(s, order="F"),
               "vel": np.zeros(s + (3,), order="F"), "prs": np.zeros(s, order="F"),
               "tmp": np.zeros(s, order="F")}
        for m in range(5):
            out["q"][I + (m,)] = self.get(f"q{m + 1}")
        out["rho"][I] = self.get("rho")
        for m, nm in enumerate("uvw"):
            out["vel"][I + (m,)] = self.get(nm)
        out["prs"][I] = self.get("prs")
        out["tmp"][I] = self.get("tmp")
        return out
