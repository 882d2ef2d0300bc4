"""Host side of the HIP solver: the reference's solver interface over the C ABI.

Method names follow the reference routines they replace, so a parity test
reads like the reference's own call sequence (src/mainloop.F90:392-476):

    solver.filterq(); solver.qswap(); solver.rhscal(); solver.rk3_update(s, dt)

Host arrays use the reference layout: Fortran order, i fastest, halo hm=5 on
every axis, variable index last, e.g. q has shape (im+11, jm+11, km+11, 5).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import AstrParams, check, dptr

HM = 5
UDF_RIEMANN2D = 1  # ASTR_UDF_RIEMANN2D: udf_bc of the 2-D Riemann example


def default_params(n: int | tuple[int, int, int], **over) -> AstrParams:
    """Taylor-Green deck values (examples/Taylor_Green_Vortex/datin/input.tgv)."""
    lib = _lib.load()
    p = AstrParams()
    if isinstance(n, int):
        n = (n, n, n)
    lib.astr_default_params(ctypes.byref(p), int(n[0]))
    p.im, p.jm, p.km = (int(v) for v in n)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def haloed_shape(p: AstrParams, extra: tuple = ()) -> tuple:
    return (p.im + 1 + 2 * HM, p.jm + 1 + 2 * HM, p.km + 1 + 2 * HM) + tuple(extra)


class AstrSolver:
    """One periodic box resident on one GPU (one astr_ctx)."""

    def __init__(self, params: AstrParams, device: int = 0):
        self.lib = _lib.load()
        self.p = params
        ctx = ctypes.c_void_p()
        check(self.lib.astr_create(ctypes.byref(params), int(device), ctypes.byref(ctx)),
              "astr_create")
        self.ctx = ctx

    def close(self) -> None:
        if self.ctx:
            self.lib.astr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- data
    def _f(self, a, shape):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return np.asfortranarray(a)

    def upload_geom(self, jacob, dxi) -> None:
        jacob = self._f(jacob, haloed_shape(self.p))
        dxi = self._f(dxi, haloed_shape(self.p, (3, 3)))
        check(self.lib.astr_upload_geom(self.ctx, dptr(jacob), dptr(dxi)), "astr_upload_geom")

    def upload_state(self, q, rho, vel, prs, tmp) -> None:
        s = haloed_shape(self.p)
        arrs = [self._f(q, s + (5,)), self._f(rho, s), self._f(vel, s + (3,)),
                self._f(prs, s), self._f(tmp, s)]
        check(self.lib.astr_upload_state(self.ctx, *[dptr(a) for a in arrs]),
              "astr_upload_state")

    def download_state(self) -> dict:
        s = haloed_shape(self.p)
        out = {"q": np.zeros(s + (5,), order="F"), "rho": np.zeros(s, order="F"),
               "vel": np.zeros(s + (3,), order="F"), "prs": np.zeros(s, order="F"),
               "tmp": np.zeros(s, order="F")}
        check(self.lib.astr_download_state(self.ctx, dptr(out["q"]), dptr(out["rho"]),
                                           dptr(out["vel"]), dptr(out["prs"]),
                                           dptr(out["tmp"])), "astr_download_state")
        return out

    def download_rhs(self) -> dict:
        s = (self.p.im + 1, self.p.jm + 1, self.p.km + 1)
        out = {"qrhs": np.zeros(s + (5,), order="F"), "dvel": np.zeros(s + (3, 3), order="F"),
               "dtmp": np.zeros(s + (3,), order="F")}
        check(self.lib.astr_download_rhs(self.ctx, dptr(out["qrhs"]), dptr(out["dvel"]),
                                         dptr(out["dtmp"])), "astr_download_rhs")
        return out

    def download_shock(self) -> np.ndarray:
        """lshock(0:im,0:jm,0:km) of the last shock sensor, as booleans."""
        out = np.zeros((self.p.im + 1, self.p.jm + 1, self.p.km + 1), dtype=np.uint8, order="F")
        check(self.lib.astr_download_shock(
            self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "astr_download_shock")
        return out.astype(bool)

    # ------------------------------------------------------- stage pieces
    def filterq(self) -> None:
        check(self.lib.astr_filterq(self.ctx), "astr_filterq")

    def qswap(self) -> None:
        check(self.lib.astr_qswap(self.ctx), "astr_qswap")

    def rhscal(self) -> None:
        check(self.lib.astr_rhscal(self.ctx), "astr_rhscal")

    def rk3_update(self, stage: int, dt: float) -> None:
        check(self.lib.astr_rk3_update(self.ctx, int(stage), float(dt)), "astr_rk3_update")

    def statcal(self) -> tuple[float, float, float]:
        out = np.zeros(3)
        check(self.lib.astr_stats(self.ctx, dptr(out)), "astr_stats")
        return float(out[0]), float(out[1]), float(out[2])

    def qmax(self) -> tuple:
        """max_q of maxmincal: the largest q(m) over the nodes 0..n
        (astr_qmax)."""
        out = np.zeros(5)
        check(self.lib.astr_qmax(self.ctx, dptr(out)), "astr_qmax")
        return tuple(float(v) for v in out)

    # ---------------------------------------------------------- steps
    def time_integration_rk(self, dt: float, stats: bool = False):
        out = np.zeros(3) if stats else None
        check(self.lib.astr_rk3_step(self.ctx, float(dt), dptr(out)), "astr_rk3_step")
        return None if out is None else tuple(float(v) for v in out)

    def steploop(self, nsteps: int, dt: float, stats: bool = False):
        out = np.zeros((nsteps, 3)) if stats else None
        check(self.lib.astr_rk3_steps(self.ctx, int(nsteps), float(dt), dptr(out)),
              "astr_rk3_steps")
        return out

    def set_spacing(self, dx: float, dy: float, dz: float) -> None:
        check(self.lib.astr_set_spacing(self.ctx, float(dx), float(dy), float(dz)),
              "astr_set_spacing")

    def line_adjoint(self, op: int, direction: int, gbar, wrap: bool = False) -> np.ndarray:
        """The transpose of the compact derivative (op 0) or filter (op 1)
        along direction (1..3) applied to gbar (haloed; nodes 0..n used):
        the haloed adjoint of the operator's input (astr_line_adjoint)."""
        g = self._f(gbar, haloed_shape(self.p))
        out = np.zeros(haloed_shape(self.p), order="F")
        check(self.lib.astr_line_adjoint(self.ctx, int(op), int(direction), int(bool(wrap)),
                                         dptr(g), dptr(out)), "astr_line_adjoint")
        return out

    def adjoint_enstrophy(self, nsteps: int, dt: float, ckpt: int = 8):
        """(J, qbar, pbar, ms): the enstrophy after nsteps RK3 steps from the
        current state and its gradient with respect to that state
        (astr_adjoint_enstrophy); ms = (forward, reversal) milliseconds."""
        J = np.zeros(1)
        qb = np.zeros(haloed_shape(self.p, (5,)), order="F")
        pb = np.zeros(haloed_shape(self.p, (6,)), order="F")
        ms = np.zeros(2)
        check(self.lib.astr_adjoint_enstrophy(self.ctx, int(nsteps), float(dt), int(ckpt),
                                              dptr(J), dptr(qb), dptr(pb), dptr(ms)),
              "astr_adjoint_enstrophy")
        return float(J[0]), qb, pb, (float(ms[0]), float(ms[1]))

    def adjoint_record(self, nsteps: int, dt: float, ckpt: int = 8) -> float:
        """The forward half of adjoint_enstrophy: nsteps RK3 steps from the
        current state, kept every ckpt steps (the last segment's stages
        stored), and J, the enstrophy of the final state
        (astr_adjoint_record)."""
        J = np.zeros(1)
        check(self.lib.astr_adjoint_record(self.ctx, int(nsteps), float(dt), int(ckpt), dptr(J)),
              "astr_adjoint_record")
        return float(J[0])

    def adjoint_reverse(self):
        """(qbar, pbar, ms): the gradient of the recorded J with respect to
        the recorded initial state, and the milliseconds of the reversal
        (astr_adjoint_reverse)."""
        qb = np.zeros(haloed_shape(self.p, (5,)), order="F")
        pb = np.zeros(haloed_shape(self.p, (6,)), order="F")
        ms = np.zeros(1)
        check(self.lib.astr_adjoint_reverse(self.ctx, dptr(qb), dptr(pb), dptr(ms)),
              "astr_adjoint_reverse")
        return qb, pb, float(ms[0])

    def set_bc(self, bctype=(1, 1, 1, 1, 1, 1), udf: int = 0) -> None:
        """boucon's faces (i=0, i=im, j=0, j=jm, k=0, k=km): 1 nothing done,
        0 the user condition udf (UDF_RIEMANN2D: udf_bc of the 2-D Riemann
        example), 23 gcnscbc (acting on faces 2 and 4), 50 zeroextrap and 60
        symmetry (any face), 41 and 42 noslip walls (isothermal at set_twall's
        temperature, adiabatic; faces 1-4), 21 outflow (faces 2 and 4), 31
        fixbc, 421 and 411 the adiabatic and isothermal slip walls (faces 3
        and 4; 411 slips where x <= set_xslip's point); none of these with a
        user condition."""
        arr = (ctypes.c_int32 * 6)(*[int(b) for b in bctype])
        check(self.lib.astr_set_bc(self.ctx, arr, int(udf)), "astr_set_bc")

    def set_twall(self, twall) -> None:
        """Wall temperature of each face for bctype 41 (noslip)."""
        arr = (ctypes.c_double * 6)(*[float(t) for t in twall])
        check(self.lib.astr_set_twall(self.ctx, arr), "astr_set_twall")

    def set_xslip(self, xslip: float) -> None:
        """The slip-nonslip point of bctype 411 (the isothermal slip wall)."""
        check(self.lib.astr_set_xslip(self.ctx, float(xslip)), "astr_set_xslip")

    def set_lengths(self, lx: float, ly: float, lz: float = 0.0) -> None:
        """gridcube's box lengths (x(i) = lx/im*i), for the boundary
        conditions that read node coordinates."""
        check(self.lib.astr_set_lengths(self.ctx, float(lx), float(ly), float(lz)),
              "astr_set_lengths")

    def boucon(self) -> None:
        check(self.lib.astr_boucon(self.ctx), "astr_boucon")

    @property
    def time(self) -> float:
        out = ctypes.c_double(0.0)
        check(self.lib.astr_get_time(self.ctx, ctypes.byref(out)), "astr_get_time")
        return out.value

    @time.setter
    def time(self, t: float) -> None:
        check(self.lib.astr_set_time(self.ctx, float(t)), "astr_set_time")

    def synchronize(self) -> None:
        check(self.lib.astr_synchronize(self.ctx), "astr_synchronize")

    def set_timing(self, on) -> None:
        """False/0 off; True/1 per kernel family (events between the
        families); 2 whole steploop calls only (two events per call, the
        device time in timing()[7])."""
        mode = int(on) if not isinstance(on, bool) else (1 if on else 0)
        check(self.lib.astr_set_timing(self.ctx, mode), "astr_set_timing")

    def timing(self) -> np.ndarray:
        """ms per family: filter, boucon+qswap, grad, gflux (central) or sensor
        (upwind steps), stats, flux, update, total, sensor (upwind rhscal)."""
        out = np.zeros(9)
        check(self.lib.astr_get_timing(self.ctx, dptr(out)), "astr_get_timing")
        return out

    def bench_copy(self, reps: int = 10) -> tuple[float, int]:
        """(ms per copy, bytes read+written per copy) of a field-sized copy."""
        ms = ctypes.c_double()
        nb = ctypes.c_int64()
        check(self.lib.astr_bench_copy(self.ctx, int(reps), ctypes.byref(ms), ctypes.byref(nb)),
              "astr_bench_copy")
        return ms.value, nb.value

    # ------------------------------------------------- field operators
    def central(self, direction: int, f) -> np.ndarray:
        """fds%central along `direction` (1..3) for every line of field f."""
        f = self._f(f, haloed_shape(self.p))
        df = np.zeros((self.p.im + 1, self.p.jm + 1, self.p.km + 1), order="F")
        check(self.lib.astr_deriv_field(self.ctx, int(direction), dptr(f), dptr(df)),
              "astr_deriv_field")
        return df

    def compact_filter(self, direction: int, f) -> np.ndarray:
        """compact_filter along `direction` (1..3) for every line of field f."""
        f = self._f(f, haloed_shape(self.p))
        ff = np.zeros((self.p.im + 1, self.p.jm + 1, self.p.km + 1), order="F")
        check(self.lib.astr_filter_field(self.ctx, int(direction), dptr(f), dptr(ff)),
              "astr_filter_field")
        return ff

    def interface_flux(self, direction: int, f, wind: int = 1, shock: bool = False) -> np.ndarray:
        """Upwind interface values of split-flux field f along `direction`:
        flux_compact (convection 2) or recons_exp (convection 1) of every
        line; haloed result, entry i the interface i+1/2 (zero where no
        interface is computed)."""
        f = self._f(f, haloed_shape(self.p))
        fh = np.zeros(haloed_shape(self.p), order="F")
        check(self.lib.astr_flux_field(self.ctx, int(direction), int(wind), int(bool(shock)),
                                       dptr(f), dptr(fh)), "astr_flux_field")
        return fh


# -------------------------------------------------------------- set-up data

class AstrBatch:
    """Contexts that advance together (astr_batch_*): one HIP graph per RK3
    step covering every member, or the line-queue engine (set_mode(2)).  The
    members stay usable on their own."""

    def __init__(self, solvers):
        self.lib = _lib.load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.ctx for s in self.solvers])
        b = ctypes.c_void_p()
        check(self.lib.astr_batch_create(arr, len(self.solvers), ctypes.byref(b)),
              "astr_batch_create")
        self.b = b

    def rk3_steps(self, nsteps: int, dt: float) -> None:
        check(self.lib.astr_batch_rk3_steps(self.b, int(nsteps), float(dt)),
              "astr_batch_rk3_steps")

    def set_mode(self, mode: int) -> None:
        """0: one HIP graph per step (default); 1: direct launches on the
        members' own streams; 2: the line-queue engine (every sweep and
        pointwise kernel one launch over all members, one graph per step)."""
        check(self.lib.astr_batch_set_mode(self.b, int(mode)), "astr_batch_set_mode")

    # the configuration bits of a queue launch (astr_batch_autotune)
    TUNE_BITS = {1: "persistent grid pulling tiles from the work queue",
                 2: "members' tiles interleaved (else in runs, largest member first)",
                 4: "one quantity per work item (else all quantities of a tile)"}
    KINDS = ("filter", "gradient", "flux")

    def autotune(self, dt: float) -> list:
        """Time every configuration of the line-queue launches for one step
        each (the members' state is restored) and keep the fastest per sweep
        kind and rows-per-lane class; returns the record of tuning()."""
        check(self.lib.astr_batch_autotune(self.b, float(dt)), "astr_batch_autotune")
        return self.tuning()

    def tuning(self) -> list:
        """The choices kept: one dict per (sweep kind, rows-per-lane class L)
        with the chosen configuration, its meaning, and the trial time of
        every configuration (microseconds per step)."""
        n = ctypes.c_int32(0)
        check(self.lib.astr_batch_get_tuning(self.b, None, 0, ctypes.byref(n)),
              "astr_batch_get_tuning")
        rows = np.zeros((max(n.value, 1), 11))
        check(self.lib.astr_batch_get_tuning(self.b, dptr(rows), n.value, ctypes.byref(n)),
              "astr_batch_get_tuning")
        out = []
        for r in rows[:n.value]:
            cfg = int(r[2])
            out.append({"kind": self.KINDS[int(r[0])], "L": int(r[1]), "config": cfg,
                        "meaning": [v for b, v in self.TUNE_BITS.items() if cfg & b],
                        "trial_us": [float(v) for v in r[3:]]})
        return out

    def set_tuning(self, record: list) -> None:
        """Apply a record of tuning() without trials."""
        choice = np.full(3 * 17, -1, dtype=np.int32)
        for e in record:
            choice[self.KINDS.index(e["kind"]) * 17 + int(e["L"])] = int(e["config"])
        check(self.lib.astr_batch_set_tuning(
            self.b, choice.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
            "astr_batch_set_tuning")

    def close(self) -> None:
        if self.b:
            self.lib.astr_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hit_state(p: AstrParams, seed: int, mt: float = 0.3, k0: float = 4.0,
              lengths=(2 * math.pi,) * 3):
    """Decaying isotropic turbulence initial field: a solenoidal random
    velocity with spectrum E(k) = A k^4 exp(-2 (k/k0)^2) built as
    div_free_gen does (src/pp.F90:2589-2837: Rogallo's construction from two
    random complex amplitudes per wavevector, Hermitian completion, inverse
    FFT; IniEnergDis, :2849-2859), from a seeded generator instead of the
    reference's unseeded random_number.  The velocity is scaled so that
    sqrt(<u.u>) / c = mt with uniform rho = 1 and T = 1 (c = 1/Ma in the
    nondimensional EoS, so Ma is taken equal to mt); p = 1/(gamma Ma^2).
    Interior nodes 0..n on each axis, node n a copy of node 0; halos zero
    (qswap fills them)."""
    rng = np.random.default_rng(seed)
    n = (p.im, p.jm, p.km)
    kx = np.fft.fftfreq(n[0], d=1.0 / n[0])
    ky = np.fft.fftfreq(n[1], d=1.0 / n[1])
    kz = np.fft.rfftfreq(n[2], d=1.0 / n[2])
    K1, K2, K3 = np.meshgrid(kx, ky, kz, indexing="ij")
    kk = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)
    k12 = np.sqrt(K1 ** 2 + K2 ** 2)
    rn = rng.random((3,) + kk.shape) * 2 * np.pi
    with np.errstate(divide="ignore", invalid="ignore"):
        ek = 2.0 / 224.7699 * kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
        amp = np.where(kk > 0, np.sqrt(ek / (4 * np.pi * kk ** 2)), 0.0)
        a1 = amp * np.exp(1j * rn[0]) * np.cos(rn[2])
        a2 = amp * np.exp(1j * rn[1]) * np.sin(rn[2])
        on_axis = k12 == 0
        u1 = np.where(on_axis, a1, (a1 * kk * K2 + a2 * K1 * K3) / (kk * k12))
        u2 = np.where(on_axis, a2, (a2 * K2 * K3 - a1 * kk * K1) / (kk * k12))
        u3 = np.where(on_axis, 0.0, -a2 * k12 / kk)
    u1[kk == 0] = u2[kk == 0] = u3[kk == 0] = 0.0
    vel = [np.fft.irfftn(u, s=n, axes=(0, 1, 2)) for u in (u1, u2, u3)]
    # solenoidal projection of the real field (the real transform's
    # Hermitian completion of the k3 = 0 and Nyquist planes breaks the
    # construction's divergence-free property there); Nyquist modes
    # removed.  The projector I - k k^T / |k|^2 and the Nyquist mask are
    # even in k, so the half spectrum of the real transform carries it.
    fk = [np.fft.rfftn(v) for v in vel]
    K = (kx[:, None, None], ky[None, :, None], kz[None, None, :])
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    k2[0, 0, 0] = 1.0
    dot = (K[0] * fk[0] + K[1] * fk[1] + K[2] * fk[2]) / k2
    nyq = np.zeros(k2.shape, dtype=bool)
    for d in range(3):
        if n[d] % 2 == 0:
            nyq |= np.abs(K[d]) == n[d] // 2
    vel = []
    for d in range(3):
        g = fk[d] - K[d] * dot
        g[nyq] = 0.0
        vel.append(np.fft.irfftn(g, s=n, axes=(0, 1, 2)))
    urms = math.sqrt(sum(float(np.mean(v * v)) for v in vel))
    vel = [v / urms for v in vel]  # sqrt(<u.u>) = 1, so Mt = Ma
    return state_from_velocity(p, np.stack(vel))


ENSEMBLE_SIZES = (64, 96, 128, 160, 192)


def ensemble_dims(nbox: int = 48, seed: int = 1, sizes=ENSEMBLE_SIZES) -> list:
    """Box sizes of the decaying-turbulence ensemble: nx, ny, nz of each of
    nbox boxes drawn uniformly from `sizes` by a generator seeded with
    `seed` (numpy's default_rng), box after box, x then y then z."""
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(sizes, 3)) for _ in range(nbox)]


def ensemble_member_seed(seed: int, member: int) -> int:
    """Seed of member `member`'s turbulence field in the ensemble of `seed`."""
    return seed * 1000 + member


def state_from_velocity(p: AstrParams, u0) -> dict:
    """The decaying-turbulence state of a velocity field u0 (3, n0, n1, n2)
    given at the unique periodic nodes: rho = 1, p = 1/(gamma Ma^2) (T = 1),
    q from fvar2q's temperature form (src/fludyna.F90:449-533); node n
    copies node 0, halos zero (qswap fills them)."""
    n = (p.im, p.jm, p.km)
    vel = [np.asarray(u0[c], dtype=np.float64) for c in range(3)]
    gamma, mach = p.gamma, p.mach
    const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach))
    const2 = gamma * (mach * mach)
    s = haloed_shape(p)
    idx = [np.mod(np.arange(n[d] + 1), n[d]) for d in range(3)]
    ix = np.ix_(*idx)
    u, v, w = (vv[ix] for vv in vel)
    rho = np.ones_like(u)
    prs = np.full_like(u, 1.0 / const2)
    tmp = prs / rho * const2
    out = {"q": np.zeros(s + (5,), order="F"), "rho": np.zeros(s, order="F"),
           "vel": np.zeros(s + (3,), order="F"), "prs": np.zeros(s, order="F"),
           "tmp": np.zeros(s, order="F")}
    I = (slice(HM, HM + n[0] + 1), slice(HM, HM + n[1] + 1), slice(HM, HM + n[2] + 1))
    out["rho"][I] = rho
    out["vel"][I + (0,)] = u
    out["vel"][I + (1,)] = v
    out["vel"][I + (2,)] = w
    out["prs"][I] = prs
    out["tmp"][I] = tmp
    out["q"][I + (0,)] = rho
    out["q"][I + (1,)] = rho * u
    out["q"][I + (2,)] = rho * v
    out["q"][I + (3,)] = rho * w
    out["q"][I + (4,)] = rho * (tmp * const1 + 0.5 * (u * u + v * v + w * w))
    return out


def box_metrics(p: AstrParams, lengths=(2 * math.pi,) * 3):
    """Metrics of the uniform box of gridcube (src/gridgeneration.F90:233-267):
    dxi = diag(1/dx, 1/dy, 1/dz), jacob = dx*dy*dz, halos included (the
    single-rank grid exchange extends the grid linearly).  An axis of 0
    cells (a two-dimensional box) has unit spacing."""
    n = (p.im, p.jm, p.km)
    dx = [lengths[d] / n[d] if n[d] else 1.0 for d in range(3)]
    s = haloed_shape(p)
    jacob = np.full(s, dx[0] * dx[1] * dx[2], order="F")
    dxi = np.zeros(s + (3, 3), order="F")
    for d in range(3):
        dxi[..., d, d] = 1.0 / dx[d]
    return jacob, dxi


def tgv_state(p: AstrParams, lengths=(2 * math.pi,) * 3):
    """Taylor-Green initial field of tgvini (src/initialisation.F90:1079-1149)
    with q from updateq/fvar2q (temperature form, src/fludyna.F90:449-533).
    Interior nodes only; halos are zero (qswap fills them)."""
    gamma, mach = p.gamma, p.mach
    const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach))
    const2 = gamma * (mach * mach)
    pinf = 1.0 / const2
    s = haloed_shape(p)
    x = (lengths[0] / p.im) * np.arange(p.im + 1)
    y = (lengths[1] / p.jm) * np.arange(p.jm + 1)
    z = (lengths[2] / p.km) * np.arange(p.km + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    rho = np.ones_like(X)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -(np.cos(X) * np.sin(Y) * np.cos(Z))
    w = np.zeros_like(X)
    prs = pinf + rho / 16.0 * 1.0 * (np.cos(2.0 * X) + np.cos(2.0 * Y)) * (np.cos(2.0 * Z) + 2.0)
    tmp = prs / rho * const2
    out = {"q": np.zeros(s + (5,), order="F"), "rho": np.zeros(s, order="F"),
           "vel": np.zeros(s + (3,), order="F"), "prs": np.zeros(s, order="F"),
           "tmp": np.zeros(s, order="F")}
    I = (slice(HM, HM + p.im + 1), slice(HM, HM + p.jm + 1), slice(HM, HM + p.km + 1))
    out["rho"][I] = rho
    out["vel"][I + (0,)] = u
    out["vel"][I + (1,)] = v
    out["vel"][I + (2,)] = w
    out["prs"][I] = prs
    out["tmp"][I] = tmp
    out["q"][I + (0,)] = rho
    out["q"][I + (1,)] = rho * u
    out["q"][I + (2,)] = rho * v
    out["q"][I + (3,)] = rho * w
    out["q"][I + (4,)] = rho * (tmp * const1 + 0.5 * (u * u + v * v + w * w))
    return out


def riemann2d_state(p: AstrParams, lengths=(1.0, 1.0), periodic_halos=False):
    """Four-quadrant 2-D Riemann initial field of riem2dini
    (src/initialisation.F90:870-921) on the unit square, split at 0.8:
    (rho, u, v, p) NE (1.5, 0, 0, 1.5), NW (0.5323, 1.206, 0, 0.30),
    SW (0.138, 1.206, 1.206, 0.029), SE (0.5323, 0, 1.206, 0.3).  The NW
    pressure literal is single precision in the reference (0.30, not
    0.30d0) and is kept as such.  T from thermal (p/rho*const2), q from
    fvar2q's temperature form (src/fludyna.F90:495-505).  With
    periodic_halos the halos (all axes) are periodic copies of nodes 0..n-1,
    otherwise the halos repeat the nearest boundary node."""
    gamma, mach = p.gamma, p.mach
    const1 = 1.0 / (gamma * (gamma - 1.0) * (mach * mach))
    const2 = gamma * (mach * mach)
    s = haloed_shape(p)
    n = (p.im, p.jm, p.km)
    idx = []
    for d in range(3):
        t = np.arange(-HM, n[d] + HM + 1)
        if n[d] == 0:
            t = np.zeros_like(t)
        elif periodic_halos:
            t = np.mod(t, n[d])
        else:
            t = np.clip(t, 0, n[d])
        idx.append(t)
    x = lengths[0] / p.im * idx[0]
    y = lengths[1] / p.jm * idx[1]
    X, Y, _ = np.meshgrid(x, y, idx[2], indexing="ij")
    ne = (X >= 0.8) & (Y >= 0.8)
    nw = (X < 0.8) & (Y >= 0.8)
    sw = (X < 0.8) & (Y < 0.8)
    se = (X >= 0.8) & (Y < 0.8)
    rho = np.select([ne, nw, sw, se], [1.5, 0.5323, 0.138, 0.5323])
    u = np.select([ne, nw, sw, se], [0.0, 1.206, 1.206, 0.0])
    v = np.select([ne, nw, sw, se], [0.0, 0.0, 1.206, 1.206])
    prs = np.select([ne, nw, sw, se], [1.5, float(np.float32(0.30)), 0.029, 0.3])
    w = np.zeros_like(rho)
    tmp = prs / rho * const2
    q = np.stack([rho, rho * u, rho * v, rho * w,
                  rho * (tmp * const1 + 0.5 * (u ** 2 + v ** 2 + w ** 2))], axis=-1)
    f = np.asfortranarray
    return {"q": f(q), "rho": f(rho), "vel": f(np.stack([u, v, w], axis=-1)), "prs": f(prs),
            "tmp": f(tmp)}


def cfl_dt(p: AstrParams, st: dict, dxi, cfl: float = 0.5) -> float:
    """The time step with the given CFL number as cflcal measures it
    (src/commcal.F90:27-87): cfl = dt * sum over the axes of the largest
    characteristic speed max(U, U - c|xi|, U + c|xi|), U = xi . u, c =
    sqrt(T)/Ma, over the nodes 0..n (the reference only reports it; its
    decks fix deltat)."""
    sl = tuple(slice(HM, HM + n + 1) for n in (p.im, p.jm, p.km))
    vel = np.asarray(st["vel"])[sl]
    css = np.sqrt(np.asarray(st["tmp"])[sl]) / p.mach
    dx = np.asarray(dxi)[sl]
    total = 0.0
    for d in range(3):
        ub = sum(dx[..., d, c] * vel[..., c] for c in range(3))
        cs = css * np.sqrt(sum(dx[..., d, c] ** 2 for c in range(3)))
        total += max(0.0, float(np.max(np.maximum(ub, np.maximum(ub - cs, ub + cs)))))
    return cfl / total


def interior(a, p: AstrParams):
    """Interior (0:im,0:jm,0:km) view of a haloed array."""
    return a[HM:HM + p.im + 1, HM:HM + p.jm + 1, HM:HM + p.km + 1]
