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
