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
        lib.orc_box_set_metrics.restype = None
        lib.orc_box_set_metrics.argtypes = [_V, _D, _D]
        lib.orc_box_set_ntype.restype = _I
        lib.orc_box_set_ntype.argtypes = [_V, _I, _I, _I]
        lib.orc_box_set_bc.restype = _I
        lib.orc_box_set_bc.argtypes = [_V, ctypes.POINTER(ctypes.c_int)]
        lib.orc_box_set_twall.restype = None
        lib.orc_box_set_twall.argtypes = [_V, ctypes.POINTER(ctypes.c_double)]
        lib.orc_box_set_xslip.restype = None
        lib.orc_box_set_xslip.argtypes = [_V, ctypes.c_double]
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
    """Box of the C restatement (periodic, or on the main path bounded on the
    axes whose npdc is 1, 2 or 4); variant 0 = main path, 1 = mini-app,
    2 = the mini-app's stage order with the main path's metric form and J
    factors."""

    def __init__(self, n, variant=0, reynolds=1600.0, mach=0.1, prandtl=0.72, gamma=1.4,
                 ref_tem=273.15, alfa=0.49, lengths=(2 * math.pi,) * 3, npdc=(3, 3, 3)):
        if isinstance(n, int):
            n = (n, n, n)
        self.n = tuple(int(v) for v in n)
        self.lib = oracle_lib()
        self.h = self.lib.orc_box_new(self.n[0], self.n[1], self.n[2], variant, reynolds, mach,
                                      prandtl, gamma, ref_tem, alfa, *lengths)
        if tuple(npdc) != (3, 3, 3) and self.lib.orc_box_set_ntype(self.h, *npdc) != 0:
            raise ValueError(f"closures {npdc} need the main variant")

    def close(self):
        if self.h:
            self.lib.orc_box_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_twall(self, twall):
        """Wall temperature of each face for bctype 41 (default 1)."""
        self.lib.orc_box_set_twall(self.h, (ctypes.c_double * 6)(*[float(t) for t in twall]))

    def set_xslip(self, xslip):
        """Slip-nonslip point of bctype 411 (default 0)."""
        self.lib.orc_box_set_xslip(self.h, float(xslip))

    def set_bc(self, bctype):
        """Boundary condition of faces 1..6 for the steps, applied in face
        order before each stage's qswap: 0 (nothing done), 23 (gcnscbc,
        acting on faces 2 and 4), 50 (zeroextrap), 60 (symmetry), 41 or 42
        (noslip isothermal or adiabatic, faces 1-4); the main variant only."""
        arr = (ctypes.c_int * 6)(*[int(v) for v in bctype])
        if self.lib.orc_box_set_bc(self.h, arr) != 0:
            raise ValueError(f"bctype {tuple(bctype)}: 0, 23, 50, 60 (41, 42 on faces 1-4) per "
                             f"face, main variant")

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

    def set_metrics(self, jacob, dxi):
        """Haloed jacob (ni, nj, nk) and dxi (ni, nj, nk, 3, 3), dxi[..., d, e] =
        d(xi_d)/d(x_e)."""
        jac = np.asfortranarray(jacob, dtype=np.float64)
        dx = np.asfortranarray(dxi, dtype=np.float64)
        self.lib.orc_box_set_metrics(self.h, _p(jac), _p(dx))

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


# ----------------------------------------------------- explicit upwind path

class UwdCfg(ctypes.Structure):
    """Mirror of orc_uwd (oracle.h)."""
    _fields_ = [("im", ctypes.c_int), ("jm", ctypes.c_int), ("km", ctypes.c_int),
                ("npdc", ctypes.c_int * 3), ("recon_schem", ctypes.c_int),
                ("lchardecomp", ctypes.c_int), ("bfacmpld", ctypes.c_double),
                ("shkcrt", ctypes.c_double), ("gamma", ctypes.c_double),
                ("mach", ctypes.c_double), ("lx", ctypes.c_double), ("ly", ctypes.c_double),
                ("lz", ctypes.c_double), ("twall", ctypes.c_double * 6),
                ("deltat", ctypes.c_double), ("xslip", ctypes.c_double)]


_uwd_declared = False


def uwd_lib():
    global _uwd_declared
    lib = oracle_lib()
    if not _uwd_declared:
        for n in ("orc_suw3", "orc_suw5", "orc_suw7", "orc_mp7", "orc_weno5", "orc_weno5z",
                  "orc_weno7", "orc_weno7z", "orc_round"):
            getattr(lib, n).restype = _F
            getattr(lib, n).argtypes = [_D]
        lib.orc_mp5.restype = _F
        lib.orc_mp5.argtypes = [_D, _D, _I]
        lib.orc_mp5ld.restype = _F
        lib.orc_mp5ld.argtypes = [_D, _F]
        lib.orc_mp7ld.restype = _F
        lib.orc_mp7ld.argtypes = [_D, _F, _I]
        lib.orc_minmod2.restype = _F
        lib.orc_minmod2.argtypes = [_F, _F]
        lib.orc_minmod4.restype = _F
        lib.orc_minmod4.argtypes = [_F, _F, _F, _F]
        lib.orc_recons_exp.restype = _F
        lib.orc_recons_exp.argtypes = [_D, _I, _I, _I, _I, _I, _F]
        lib.orc_sw_split.argtypes = [_F, _F, _F, _D, _F, _F, _D, _D, _F, _D, _D]
        lib.orc_chardecomp.restype = _I
        lib.orc_chardecomp.argtypes = [_F, _F, _F, _F, _F, _D, _D, _F, _F, _F, _F, _D, _D, _D,
                                       _D]
        lib.orc_iwind8.restype = _I
        lib.orc_iwind8.argtypes = [_I, _I, _I, _I, _I]
        P = ctypes.POINTER(UwdCfg)
        U8 = ctypes.POINTER(ctypes.c_ubyte)
        lib.orc_gradcal.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_ducros.argtypes = [P, _D, _D, _D, U8]
        lib.orc_convrsduwd.restype = _I
        lib.orc_convrsduwd.argtypes = [P, _D, _D, _D, _D, _D, _D, _D, U8, _D]
        lib.orc_uwd_rhscal.restype = _I
        lib.orc_uwd_rhscal.argtypes = [P, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, U8]
        lib.orc_cmp_rhscal.restype = _I
        lib.orc_cmp_rhscal.argtypes = [P, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, U8]
        lib.orc_flux_compact.argtypes = [_I, _I, _I, _F, _D, _D]
        lib.orc_flux_compact_init.restype = _I
        lib.orc_flux_compact_init.argtypes = [_I, _I, _I, _F, ctypes.POINTER(_I),
                                              ctypes.POINTER(_I), _D]
        lib.orc_mplimiter.restype = _F
        lib.orc_mplimiter.argtypes = [_D, _F, _I, _I, _I, _I]
        I6 = ctypes.POINTER(ctypes.c_int)
        lib.orc_riem2d_shockspeed.restype = _F
        lib.orc_riem2d_shockspeed.argtypes = [_F, _F, _F, _F, _F]
        lib.orc_riem2d_bc.argtypes = [P, I6, _F, _D, _D, _D, _D, _D]
        lib.orc_uwd_qswap.argtypes = [P, _D, _D, _D, _D, _D]
        lib.orc_gcnscbc.restype = None
        lib.orc_gcnscbc.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_zeroextrap.restype = None
        lib.orc_zeroextrap.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_symmetry.restype = None
        lib.orc_symmetry.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_wall.restype = None
        lib.orc_wall.argtypes = [P, _I, _I, _D, _D, _D, _D, _D]
        lib.orc_outflow.restype = None
        lib.orc_outflow.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_fixbc.restype = None
        lib.orc_fixbc.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_slipisot.restype = None
        lib.orc_slipisot.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_slipadib.restype = None
        lib.orc_slipadib.argtypes = [P, _I, _D, _D, _D, _D, _D]
        lib.orc_uwd_rk3_step.restype = _I
        lib.orc_uwd_rk3_step.argtypes = [P, _I, _I, I6, _I, _F, _F, _D, _D, _D, _D, _D, _D, _D]
        _uwd_declared = True
    return lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def recon(name: str, u, *extra) -> float:
    """One scalar reconstruction of src/flux.F90 (suw3/5/7, mp5, mp7, mp5ld,
    mp7ld, weno5/5z/7/7z, round)."""
    lib = uwd_lib()
    u = _c(u)
    fn = getattr(lib, "orc_" + name)
    if name == "mp5":
        return fn(_p(u), None, -1)
    return fn(_p(u), *extra)


def recons_exp(f, inode, dim, ntype, reschem, shock=True, bfacmpld=0.7) -> float:
    f = _c(f)
    return uwd_lib().orc_recons_exp(_p(f), inode, dim, ntype, reschem, int(shock), bfacmpld)


def flux_compact(ntype, f, plus=True, bfac=0.7) -> np.ndarray:
    """fh(-1:dim) of flux_compact for a line f(-hm:dim+hm)."""
    f = _c(f)
    dim = len(f) - 1 - 2 * HM
    fh = np.zeros(dim + 2)
    uwd_lib().orc_flux_compact(ntype, dim, int(plus), bfac, _p(f), _p(fh))
    return fh


def mplimiter(f5, fl, shock, inode, dim, ntype) -> float:
    return uwd_lib().orc_mplimiter(_p(_c(f5)), fl, int(shock), inode, dim, ntype)


def sw_split(rho, vel, prs, tmp, q, dxi, jacob, gamma=1.4, mach=0.1):
    """(F+, F-) of flux_steger_warming at one node."""
    fp = np.zeros(5)
    fm = np.zeros(5)
    uwd_lib().orc_sw_split(gamma, mach, rho, _p(_c(vel)), prs, tmp, _p(_c(q)), _p(_c(dxi)),
                           jacob, _p(fp), _p(fm))
    return fp, fm


def chardecomp(left, right, gamma=1.4):
    """REV, LEV (5x5, REV[m-1, n-1] = REV(m, n)) of chardecomp from two
    states given as dicts with tmp, rho, prs, E, vel(3), ddi(3)."""
    R = np.zeros(25)
    L = np.zeros(25)
    rc = uwd_lib().orc_chardecomp(gamma, left["tmp"], left["rho"], left["prs"], left["E"],
                                  _p(_c(left["vel"])), _p(_c(left["ddi"])), right["tmp"],
                                  right["rho"], right["prs"], right["E"], _p(_c(right["vel"])),
                                  _p(_c(right["ddi"])), _p(R), _p(L))
    if rc:
        raise ValueError("zero metric normal")
    return R.reshape(5, 5), L.reshape(5, 5)


class UwdBox:
    """The explicit upwind right-hand side on one haloed box, host layout
    (Fortran order, halos hm = 5 on every axis; km = 0 is a 2-D box)."""

    def __init__(self, dims, npdc=(3, 3, 3), recon_schem=3, lchardecomp=True, bfacmpld=0.7,
                 shkcrt=0.01, gamma=1.4, mach=1.0, lengths=(1.0, 1.0, 0.0),
                 twall=(1.0,) * 6, xslip=0.0):
        im, jm, km = dims
        self.cfg = UwdCfg(im, jm, km, (ctypes.c_int * 3)(*npdc), recon_schem, int(lchardecomp),
                          bfacmpld, shkcrt, gamma, mach, *lengths,
                          (ctypes.c_double * 6)(*[float(t) for t in twall]), 0.0,
                          float(xslip))
        self.dims = (im, jm, km)
        self.hshape = (im + 1 + 2 * HM, jm + 1 + 2 * HM, km + 1 + 2 * HM)
        self.ishape = (im + 1, jm + 1, km + 1)

    def _h(self, a, n=None):
        shape = self.hshape + ((n,) if n else ())
        a = np.asarray(a, dtype=np.float64)
        assert a.shape == shape, (a.shape, shape)
        return np.asfortranarray(a)

    def rhscal(self, st, dxi, jacob, explicit_fd=True, compact=False):
        """qrhs, dvel, dtmp, ssf, lshock of the upwind rhscal for state st
        (dict q, rho, vel, prs, tmp; haloed): convrsduwd, or convrsdcmp
        with compact=True."""
        lib = uwd_lib()
        q = self._h(st["q"], 5)
        vel = self._h(st["vel"], 3)
        rho, prs, tmp = (self._h(st[k]) for k in ("rho", "prs", "tmp"))
        dx = self._h(dxi, 9) if np.asarray(dxi).ndim == 4 else np.asfortranarray(
            np.asarray(dxi).reshape(self.hshape + (9,), order="F"))
        jac = self._h(jacob)
        qrhs = np.zeros(self.ishape + (5,), order="F")
        dvel = np.zeros(self.ishape + (9,), order="F")
        dtmp = np.zeros(self.ishape + (3,), order="F")
        ssf = np.zeros(self.hshape, order="F")
        lsh = np.zeros(self.ishape, dtype=np.uint8, order="F")
        fn = lib.orc_cmp_rhscal if compact else lib.orc_uwd_rhscal
        rc = fn(ctypes.byref(self.cfg), int(explicit_fd), _p(q), _p(rho),
                                _p(vel), _p(prs), _p(tmp), _p(dx), _p(jac), _p(qrhs),
                                _p(dvel), _p(dtmp), _p(ssf),
                                lsh.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        if rc:
            raise ValueError(f"orc_uwd_rhscal: unsupported configuration ({rc})")
        return {"qrhs": qrhs, "dvel": dvel.reshape(self.ishape + (3, 3), order="F"),
                "dtmp": dtmp, "ssf": ssf, "lshock": lsh.astype(bool)}

    def convrsduwd(self, st, dxi, jacob, lshock=None):
        """qrhs accumulated by convrsduwd alone (from zero, not negated)."""
        lib = uwd_lib()
        q = self._h(st["q"], 5)
        vel = self._h(st["vel"], 3)
        rho, prs, tmp = (self._h(st[k]) for k in ("rho", "prs", "tmp"))
        dx = np.asfortranarray(np.asarray(dxi, dtype=np.float64).reshape(self.hshape + (9,),
                                                                         order="F"))
        jac = self._h(jacob)
        qrhs = np.zeros(self.ishape + (5,), order="F")
        ls = None
        if lshock is not None:
            lsa = np.asfortranarray(np.asarray(lshock, dtype=np.uint8))
            ls = lsa.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        rc = lib.orc_convrsduwd(ctypes.byref(self.cfg), _p(q), _p(rho), _p(vel), _p(prs),
                                _p(tmp), _p(dx), _p(jac), ls, _p(qrhs))
        if rc:
            raise ValueError(f"orc_convrsduwd: unsupported configuration ({rc})")
        return qrhs

    def _state(self, st):
        """Fortran-ordered float64 copies of a state dict (q, rho, vel, prs,
        tmp), which the in-place routines below update."""
        return {"q": self._h(st["q"], 5).copy(order="F"), "vel": self._h(st["vel"], 3).copy(order="F"),
                **{k: self._h(st[k]).copy(order="F") for k in ("rho", "prs", "tmp")}}

    @staticmethod
    def _faces(bctype):
        return (ctypes.c_int * 6)(*[int(b) for b in bctype])

    def riem2d_bc(self, st, time, bctype=(0, 0, 0, 0, 1, 1)):
        """A copy of st with the 2-D Riemann example's udf_bc applied on the
        faces whose bctype is 0, at the given time."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_riem2d_bc(ctypes.byref(self.cfg), self._faces(bctype), float(time), _p(o["q"]),
                          _p(o["rho"]), _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def gcnscbc(self, st, face):
        """A copy of st after gcnscbc (bctype 23) on face 2 or 4."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_gcnscbc(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]), _p(o["vel"]),
                        _p(o["prs"]), _p(o["tmp"]))
        return o

    def zeroextrap(self, st, face):
        """A copy of st after zeroextrap (bctype 50) on face 1..6."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_zeroextrap(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                           _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def outflow(self, st, face, dt):
        """A copy of st after outflow (bctype 21) on face 2 or 4, with the
        step's time step dt."""
        lib = uwd_lib()
        o = self._state(st)
        self.cfg.deltat = float(dt)
        lib.orc_outflow(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                        _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def slipisot(self, st, face):
        """A copy of st after slipisotwall (bctype 411) on face 3 or 4."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_slipisot(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                         _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def slipadib(self, st, face):
        """A copy of st after slipadibwall (bctype 421) on face 3 or 4."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_slipadib(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                         _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def fixbc(self, st, face):
        """A copy of st after fixbc (bctype 31) on face 1..6."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_fixbc(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                      _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def wall(self, st, face, isothermal):
        """A copy of st after noslip (bctype 41, isothermal, at the face's
        twall) or noslip_adibatic (42) on face 1..4."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_wall(ctypes.byref(self.cfg), int(face), int(bool(isothermal)), _p(o["q"]),
                     _p(o["rho"]), _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def symmetry(self, st, face):
        """A copy of st after symmetry (bctype 60) on face 1..6."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_symmetry(ctypes.byref(self.cfg), int(face), _p(o["q"]), _p(o["rho"]),
                         _p(o["vel"]), _p(o["prs"]), _p(o["tmp"]))
        return o

    def qswap(self, st):
        """A copy of st after the single-rank qswap."""
        lib = uwd_lib()
        o = self._state(st)
        lib.orc_uwd_qswap(ctypes.byref(self.cfg), _p(o["q"]), _p(o["rho"]), _p(o["vel"]),
                          _p(o["prs"]), _p(o["tmp"]))
        return o

    def rk3_steps(self, st, dxi, jacob, dt, nsteps=1, time=0.0, bctype=(1, 1, 1, 1, 1, 1), udf=0,
                  compact=False, explicit_fd=True):
        """A copy of st advanced nsteps RK3 steps of the upwind path from the
        given start time (boundary conditions by bctype/udf)."""
        lib = uwd_lib()
        o = self._state(st)
        dx = np.asfortranarray(np.asarray(dxi, dtype=np.float64).reshape(self.hshape + (9,),
                                                                         order="F"))
        jac = self._h(jacob)
        t = float(time)
        for _ in range(nsteps):
            rc = lib.orc_uwd_rk3_step(ctypes.byref(self.cfg), int(compact), int(explicit_fd),
                                      self._faces(bctype), int(udf), t, float(dt), _p(dx), _p(jac),
                                      _p(o["q"]), _p(o["rho"]), _p(o["vel"]), _p(o["prs"]),
                                      _p(o["tmp"]))
            if rc:
                raise ValueError(f"orc_uwd_rk3_step: unsupported configuration ({rc})")
            t = t + float(dt)
        return o

    def ducros(self, dvel, prs):
        lib = uwd_lib()
        dv = np.asfortranarray(np.asarray(dvel, dtype=np.float64).reshape(self.ishape + (9,),
                                                                          order="F"))
        p = self._h(prs)
        ssf = np.zeros(self.hshape, order="F")
        lsh = np.zeros(self.ishape, dtype=np.uint8, order="F")
        lib.orc_ducros(ctypes.byref(self.cfg), _p(dv), _p(p), _p(ssf),
                       lsh.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return ssf, lsh.astype(bool)


def prim_to_state(rho, u, v, w, p, gamma=1.4, mach=1.0):
    """Conservative q and T from primitives (nondimensional EoS of
    src/fludyna.F90:449-533: const2 = gamma Ma^2, const6 = 1/(gamma-1))."""
    const2 = gamma * mach * mach
    const6 = 1.0 / (gamma - 1.0)
    q = np.stack([rho, rho * u, rho * v, rho * w,
                  p * const6 + 0.5 * rho * (u * u + v * v + w * w)], axis=-1)
    tmp = p / rho * const2
    return {"q": np.asfortranarray(q), "rho": np.asfortranarray(rho),
            "vel": np.asfortranarray(np.stack([u, v, w], axis=-1)),
            "prs": np.asfortranarray(p), "tmp": np.asfortranarray(tmp)}
