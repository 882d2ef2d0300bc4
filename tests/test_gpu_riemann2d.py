"""The 2-D Riemann problem on the GPU (through the C ABI) against the
checker: the example's boundary condition, the 2-D qswap and whole upwind
RK3 steps on boxes bounded at both ends (ntype 4 on i and j).

Every piece runs without FMA contraction in the reference's operation
order on both sides, so the fields are expected to agree bit for bit; the
assertions allow 1e-13 of each field's magnitude and every test prints
whether the fields were bit-equal.
"""
import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle as po

pytestmark = pytest.mark.gpu
HM = 5
FACES = (0, 0, 0, 0, 1, 1)


def nerr(a, b):
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def params(dims, npdc=(4, 4, 3), recon=3, lchar=True, compact=False):
    return ag.default_params(dims, npdci=npdc[0], npdcj=npdc[1], npdck=npdc[2],
                             convection=2 if compact else 1, deriv_explicit=1,
                             recon_schem=recon, lchardecomp=int(lchar), lfilter=0, diffterm=0,
                             shkcrt=0.01, mach=1.0)


def setup(dims, npdc=(4, 4, 3), recon=3, lchar=True, compact=False, perturb=True, seed=0):
    p = params(dims, npdc, recon, lchar, compact)
    st = ag.riemann2d_state(p)
    if perturb:
        # a smooth perturbation so that the extrapolated faces see slopes
        rng = np.random.default_rng(seed)
        shape = st["rho"].shape
        x = np.linspace(0, 1, shape[0])[:, None, None]
        y = np.linspace(0, 1, shape[1])[None, :, None]
        f = 1.0 + 0.02 * np.sin(2 * np.pi * (x + 2 * y) + rng.random())
        prim = {"rho": st["rho"] * f, "prs": st["prs"] * (2.0 - f)}
        vel = st["vel"] * f[..., None]
        st = po.prim_to_state(prim["rho"], vel[..., 0], vel[..., 1], vel[..., 2], prim["prs"],
                              mach=p.mach)
    s = ag.AstrSolver(p)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s.upload_geom(jac, dxi)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    box = po.UwdBox(dims, npdc=npdc, recon_schem=recon, lchardecomp=lchar, bfacmpld=0.7,
                    shkcrt=0.01, gamma=p.gamma, mach=p.mach)
    return s, box, st, jac, dxi.reshape(dxi.shape[:3] + (9,), order="F")


def compare_plane(g, ref, label, tol=1e-13):
    """Fields of the plane k = 0 (halos of i and j included)."""
    exact = True
    for k in ("q", "rho", "vel", "prs", "tmp"):
        a, b = g[k][:, :, HM], ref[k][:, :, HM]
        assert np.all(np.isfinite(b)), (label, k, "checker state not finite")
        e = nerr(a, b)
        exact &= bool(np.array_equal(a, b))
        assert e <= tol, (label, k, e)
    print(f"  {label}: {'bit-equal' if exact else 'not bit-equal'}", flush=True)
    return exact


def test_boucon_matches_checker():
    dims = (64, 60, 0)
    print(f"case boucon dims={dims}", flush=True)
    s, box, st, _, _ = setup(dims)
    for faces in (FACES, (1, 0, 1, 1, 1, 1), (1, 1, 1, 0, 1, 1), (0, 1, 0, 1, 1, 1)):
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        s.set_bc(faces, ag.UDF_RIEMANN2D)
        s.time = 0.0731
        s.boucon()
        g = s.download_state()
        ref = box.riem2d_bc(st, 0.0731, bctype=faces)
        compare_plane(g, ref, f"faces {faces}")
    s.close()


@pytest.mark.parametrize("npdc", [(4, 4, 3), (3, 4, 3), (4, 3, 3), (3, 3, 3)])
def test_qswap_two_dimensional_matches_checker(npdc):
    dims = (48, 40, 0)
    print(f"case qswap 2-D npdc={npdc}", flush=True)
    s, box, st, _, _ = setup(dims, npdc=npdc)
    s.qswap()
    g = s.download_state()
    ref = box.qswap(st)
    compare_plane(g, ref, "plane k=0")
    # the pressure of the planes k = -1, +1 (the one k-halo value read)
    for kk in (HM - 1, HM + 1):
        assert np.array_equal(g["prs"][HM:-HM, HM:-HM, kk], ref["prs"][HM:-HM, HM:-HM, kk])
    s.close()


CASES = [
    ((64, 60, 0), (4, 4, 3), 3, True, False),
    ((64, 60, 0), (4, 4, 3), 5, True, False),
    ((56, 64, 0), (4, 4, 3), 1, False, False),
    ((48, 40, 0), (3, 4, 3), 3, True, False),
    ((64, 60, 0), (4, 4, 3), 3, True, True),
]


@pytest.mark.parametrize("dims,npdc,recon,lchar,compact", CASES)
def test_riemann2d_steps_match_checker(dims, npdc, recon, lchar, compact):
    print(f"case steps dims={dims} npdc={npdc} recon={recon} lchar={lchar} compact={compact}",
          flush=True)
    s, box, st, jac, dxi9 = setup(dims, npdc, recon, lchar, compact)
    s.set_bc(FACES, ag.UDF_RIEMANN2D)
    dt, nsteps = 0.5 / max(dims) / 1.8, 6
    s.steploop(nsteps, dt)
    g = s.download_state()
    ref = box.rk3_steps(st, dxi9, jac, dt, nsteps=nsteps, bctype=FACES, udf=1, compact=compact)
    compare_plane(g, ref, f"{nsteps} steps")
    t = 0.0
    for _ in range(nsteps):
        t = t + dt
    assert s.time == t
    s.close()


def test_riemann2d_step_equals_its_pieces():
    dims = (48, 44, 0)
    print(f"case pieces dims={dims}", flush=True)
    s1, _, st, _, _ = setup(dims)
    s2, _, _, _, _ = setup(dims)
    for s in (s1, s2):
        s.set_bc(FACES, ag.UDF_RIEMANN2D)
        s.time = 0.01
    dt = 0.004
    for _ in range(2):
        s1.time_integration_rk(dt)
        for stage in (1, 2, 3):
            s2.boucon()
            s2.qswap()
            s2.rhscal()
            s2.rk3_update(stage, dt)
        s2.time = s2.time + dt
    a, b = s1.download_state(), s2.download_state()
    for k in a:
        assert np.array_equal(a[k][:, :, HM], b[k][:, :, HM]), k
    s1.close()
    s2.close()


def test_riemann2d_long_lines_match_checker():
    """x lines longer than one 256-interface tile and beyond the compact
    line sweeps' limit, bounded at both ends."""
    dims = (600, 40, 0)
    print(f"case long steps dims={dims}", flush=True)
    s, box, st, jac, dxi9 = setup(dims)
    s.set_bc(FACES, ag.UDF_RIEMANN2D)
    dt = 0.5 / 600 / 1.8
    s.steploop(2, dt)
    g = s.download_state()
    ref = box.rk3_steps(st, dxi9, jac, dt, nsteps=2, bctype=FACES, udf=1)
    compare_plane(g, ref, "2 steps")
    s.close()


def test_boundary_and_two_dimensional_rejections():
    dims = (32, 32, 0)
    s, _, _, _, _ = setup(dims)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.set_bc((22, 0, 0, 0, 1, 1), ag.UDF_RIEMANN2D)  # outflow NSCBC: not built
    with pytest.raises(ag.AstrError, match="ASTR_E_ARG"):
        s.set_bc(FACES, 0)
    with pytest.raises(ag.AstrError, match="ASTR_E_ARG"):
        s.set_bc(FACES, 7)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.filterq()
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.time_integration_rk(1e-3, stats=True)
    s.set_bc(FACES, ag.UDF_RIEMANN2D)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        ag.AstrBatch([s])
    s.close()


def pulse_state(dims, mach, u=0.3, v=0.2, amp=0.05, seed=0):
    """Uniform flow with a pressure pulse and a smooth perturbation, at a
    sound speed sqrt(gamma p/rho) of about 1/Ma (gcnscbc truncates it to an
    integer, which must not be 0)."""
    shape = (dims[0] + 1 + 2 * HM, dims[1] + 1 + 2 * HM, 1 + 2 * HM)
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, shape[0])[:, None, None] * np.ones(shape)
    y = np.linspace(0, 1, shape[1])[None, :, None] * np.ones(shape)
    pinf = 1.0 / (1.4 * mach * mach)
    g = amp * np.exp(-((x - 0.6) ** 2 + (y - 0.7) ** 2) / 0.02)
    f = 0.01 * np.sin(2 * np.pi * (x + 2 * y) + rng.random())
    return po.prim_to_state(1.0 + g + f, u + g, v - 0.5 * g, 0.0 * x, pinf * (1.0 + g - f),
                            mach=mach)


@pytest.mark.parametrize("v", [0.2, -1.2])
def test_gcnscbc_steps_match_checker(v):
    """gcnscbc (bctype 23) on faces 2 and 4 through whole upwind steps against
    the checker; v = -1.2 takes face 4's inflow branch (integer normal
    velocity -1)."""
    dims = (64, 60, 0)
    mach = 0.4
    print(f"case gcnscbc dims={dims} v={v}", flush=True)
    p = ag.default_params(dims, npdci=4, npdcj=4, npdck=3, convection=1, deriv_explicit=1,
                          recon_schem=3, lchardecomp=1, lfilter=0, diffterm=0, shkcrt=0.01,
                          mach=mach)
    st = pulse_state(dims, mach, v=v)
    s = ag.AstrSolver(p)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s.upload_geom(jac, dxi)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    faces = (1, 23, 1, 23, 1, 1)
    s.set_lengths(1.0, 1.0, 0.0)
    s.set_bc(faces, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), recon_schem=3, lchardecomp=True, bfacmpld=0.7,
                    shkcrt=0.01, gamma=p.gamma, mach=mach)
    dxi9 = dxi.reshape(dxi.shape[:3] + (9,), order="F")
    s.boucon()
    compare_plane(s.download_state(), _boucon_ref(box, st, faces), "boucon")
    dt = 0.2 / 64 * mach
    s.steploop(5, dt)
    g = s.download_state()
    ref = box.rk3_steps(st, dxi9, jac, dt, nsteps=5, bctype=faces, udf=0)
    compare_plane(g, ref, "5 steps")
    s.close()


def _boucon_ref(box, st, faces):
    o = st
    for f, b in enumerate(faces):
        if b == 23:
            o = box.gcnscbc(o, f + 1)
        elif b == 50:
            o = box.zeroextrap(o, f + 1)
        elif b == 60:
            o = box.symmetry(o, f + 1)
        elif b in (41, 42):
            o = box.wall(o, f + 1, b == 41)
        elif b == 21:
            o = box.outflow(o, f + 1, 0.0)
        elif b == 31:
            o = box.fixbc(o, f + 1)
        elif b == 421:
            o = box.slipadib(o, f + 1)
        elif b == 411:
            o = box.slipisot(o, f + 1)
    return o


@pytest.mark.parametrize("faces", [(50, 50, 50, 50, 1, 1), (50, 23, 50, 23, 1, 1),
                                   (60, 60, 60, 60, 1, 1), (60, 23, 50, 60, 1, 1),
                                   (41, 42, 41, 42, 1, 1), (50, 41, 60, 42, 1, 1),
                                   (50, 21, 50, 21, 1, 1), (50, 50, 31, 31, 1, 1),
                                   (50, 21, 421, 421, 1, 1), (50, 21, 411, 411, 1, 1)])
def test_zeroextrap_steps_match_checker(faces):
    """zeroextrap (bctype 50), symmetry (60), the noslip walls (41 at its
    face's wall temperature, 42), the adiabatic slip wall (421, faces 3 and
    4), the isothermal one (411, slipping up to x = 0.5), fixbc (31) and
    outflow (21, faces 2 and 4) on the faces
    of a 2-D box, alone, mixed and with gcnscbc on faces 2 and 4 (the
    reference's face order 1..6), through boucon and whole upwind steps
    against the checker (boucon alone runs with the time step of no step, 0)."""
    dims = (64, 60, 0)
    mach = 0.4
    print(f"case zeroextrap dims={dims} faces={faces}", flush=True)
    p = ag.default_params(dims, npdci=4, npdcj=4, npdck=3, convection=1, deriv_explicit=1,
                          recon_schem=3, lchardecomp=1, lfilter=0, diffterm=0, shkcrt=0.01,
                          mach=mach)
    st = pulse_state(dims, mach)
    s = ag.AstrSolver(p)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s.upload_geom(jac, dxi)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.set_lengths(1.0, 1.0, 0.0)
    tw = (1.05, 0.95, 1.1, 0.9, 1.0, 1.0)
    s.set_twall(tw)
    s.set_xslip(0.5)
    s.set_bc(faces, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), recon_schem=3, lchardecomp=True, bfacmpld=0.7,
                    shkcrt=0.01, gamma=p.gamma, mach=mach, twall=tw, xslip=0.5)
    dxi9 = dxi.reshape(dxi.shape[:3] + (9,), order="F")
    s.boucon()
    compare_plane(s.download_state(), _boucon_ref(box, st, faces), "boucon")
    dt = 0.2 / 64 * mach
    s.steploop(5, dt)
    g = s.download_state()
    ref = box.rk3_steps(st, dxi9, jac, dt, nsteps=5, bctype=faces, udf=0)
    compare_plane(g, ref, "5 steps")
    s.close()


def test_zeroextrap_rejected_where_not_built():
    p = ag.default_params((32, 32, 0), npdci=4, npdcj=4, npdck=3, convection=1,
                          deriv_explicit=1, lfilter=0, diffterm=0)
    s = ag.AstrSolver(p)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.set_bc((50, 0, 1, 1, 1, 1), ag.UDF_RIEMANN2D)  # with a user condition
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.set_bc((1, 1, 1, 1, 50, 1), 0)  # a k face of a 2-D box
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.set_bc((1, 1, 1, 1, 41, 1), 0)  # walls act on faces 1-4
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.set_bc((421, 1, 1, 1, 1, 1), 0)  # the slip wall acts on faces 3 and 4
    with pytest.raises(ag.AstrError, match="ASTR_E_ARG"):
        s.set_twall((1.0, 0.0, 1.0, 1.0, 1.0, 1.0))
    s.close()
