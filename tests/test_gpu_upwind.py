"""Parity of the HIP upwind path (through the C ABI) with the checker.

The device kernels keep the reference's expression order with FMA
contraction off, as the checker does, so qrhs, dvel and the shock flags are
expected to agree bit for bit; the assertions allow 1e-13 of each field's
magnitude (a limiter branch taken differently would show as an O(1)
relative difference at that node) and every test prints whether the
fields were bit-equal.
"""
import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle as po

pytestmark = pytest.mark.gpu
HM = 5


def quadrant_state(dims, mach=1.0, seed=0, smooth=0.02):
    """The four-quadrant Riemann states of riem2dini (src/initialisation.F90:
    870-921; (rho, u, v, p) NE (1.5,0,0,1.5), NW (0.5323,1.206,0,0.3), SW
    (0.138,1.206,1.206,0.029), SE (0.5323,0,1.206,0.3), split at 0.8 of
    the unit box), extruded in z with a small smooth perturbation so that
    every axis carries both shocks and smooth flow.  Halos included."""
    im, jm, km = dims
    rng = np.random.default_rng(seed)
    t = [np.arange(-HM, n + HM + 1) / max(n, 1) for n in dims]
    X, Y, Z = np.meshgrid(*t, indexing="ij")
    ne = (X >= 0.8) & (Y >= 0.8)
    nw = (X < 0.8) & (Y >= 0.8)
    sw = (X < 0.8) & (Y < 0.8)
    se = (X >= 0.8) & (Y < 0.8)
    rho = np.select([ne, nw, sw, se], [1.5, 0.5323, 0.138, 0.5323])
    u = np.select([ne, nw, sw, se], [0.0, 1.206, 1.206, 0.0])
    v = np.select([ne, nw, sw, se], [0.0, 0.0, 1.206, 1.206])
    p = np.select([ne, nw, sw, se], [1.5, 0.3, 0.029, 0.3])
    pert = smooth * np.sin(2 * np.pi * (X + 2 * Y + 3 * Z) + rng.random())
    rho = rho * (1 + pert)
    w = 0.1 * np.cos(2 * np.pi * Z) * (km > 0)
    return po.prim_to_state(rho, u + 0.05 * pert, v, w, p, mach=mach)


def pair(dims, npdc=(3, 3, 3), recon=3, lchar=True, shkcrt=0.01, explicit=True, compact=False):
    p = ag.default_params(dims, npdci=npdc[0], npdcj=npdc[1], npdck=npdc[2],
                          convection=2 if compact else 1,
                          deriv_explicit=int(explicit), recon_schem=recon,
                          lchardecomp=int(lchar), lfilter=0, diffterm=0, shkcrt=shkcrt,
                          mach=1.0)
    s = ag.AstrSolver(p)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s.upload_geom(jac, dxi)
    box = po.UwdBox(dims, npdc=npdc, recon_schem=recon, lchardecomp=lchar, bfacmpld=0.7,
                    shkcrt=shkcrt, gamma=p.gamma, mach=p.mach)
    return s, box, jac, dxi.reshape(dxi.shape[:3] + (9,), order="F")


def run_both(dims, npdc, recon, lchar, explicit=True, shkcrt=0.01, seed=0, compact=False):
    s, box, jac, dxi9 = pair(dims, npdc, recon, lchar, shkcrt, explicit, compact)
    st = quadrant_state(dims, seed=seed)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.rhscal()
    g = s.download_rhs()
    ref = box.rhscal(st, dxi9, jac, explicit_fd=explicit, compact=compact)
    sensor = lchar if compact else (recon == 5 or lchar)
    lsh = s.download_shock() if sensor else None
    s.close()
    return g, ref, lsh


def nerr(a, b):
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


CASES = [
    ((40, 36, 32), (3, 3, 3), 3, True),
    ((40, 36, 32), (3, 3, 3), 0, False),
    ((40, 36, 32), (1, 2, 3), 3, True),
    ((40, 36, 32), (4, 1, 2), 5, True),
    ((36, 40, 32), (2, 3, 1), 1, True),
    ((36, 34, 32), (3, 2, 3), 2, False),
    ((36, 34, 32), (1, 3, 3), 6, True),
    ((36, 34, 32), (3, 3, 3), -1, False),
]


@pytest.mark.parametrize("dims,npdc,recon,lchar", CASES)
def test_upwind_rhs_matches_checker(dims, npdc, recon, lchar):
    print(f"case dims={dims} npdc={npdc} recon={recon} lchar={lchar}", flush=True)
    g, ref, lsh = run_both(dims, npdc, recon, lchar)
    if lsh is not None:
        nflag = int(ref["lshock"].sum())
        print(f"  shock nodes {nflag} of {ref['lshock'].size}")
        assert np.array_equal(lsh, ref["lshock"])
    for m in range(3):
        for d in range(3):
            e = nerr(g["dvel"][..., m, d], ref["dvel"][..., m, d])
            assert e <= 1e-13, ("dvel", m, d, e)
    exact = True
    for m in range(5):
        a, b = g["qrhs"][..., m], ref["qrhs"][..., m]
        e = nerr(a, b)
        exact &= bool(np.array_equal(a, b))
        print(f"  qrhs{m + 1}: {e:.2e}")
        assert e <= 1e-13, (m, e)
    print("  bit-equal" if exact else "  not bit-equal")


CMP_CASES = [
    ((40, 36, 32), (3, 3, 3), True),
    ((40, 36, 32), (1, 2, 4), True),
    ((36, 34, 32), (4, 4, 3), False),
    ((36, 34, 32), (2, 1, 3), True),
]


@pytest.mark.parametrize("dims,npdc,lchar", CMP_CASES)
def test_compact_upwind_rhs_matches_checker(dims, npdc, lchar):
    """convrsdcmp ('543c'): compact interface fluxes solved line by line with
    the sequential Thomas recurrence, MP limiter about them."""
    print(f"case compact dims={dims} npdc={npdc} lchar={lchar}", flush=True)
    g, ref, lsh = run_both(dims, npdc, 3, lchar, compact=True)
    if lsh is not None:
        assert np.array_equal(lsh, ref["lshock"])
    exact = True
    for m in range(5):
        a, b = g["qrhs"][..., m], ref["qrhs"][..., m]
        e = nerr(a, b)
        exact &= bool(np.array_equal(a, b))
        print(f"  qrhs{m + 1}: {e:.2e}")
        assert e <= 1e-13, (m, e)
    print("  bit-equal" if exact else "  not bit-equal")


def test_compact_upwind_two_dimensional():
    dims = (64, 60, 0)
    g, ref, lsh = run_both(dims, (4, 2, 3), 3, True, compact=True)
    assert np.array_equal(lsh, ref["lshock"])
    for m in range(5):
        assert nerr(g["qrhs"][..., m], ref["qrhs"][..., m]) <= 1e-13


def test_upwind_rhs_two_dimensional_quadrants():
    dims = (64, 60, 0)
    print(f"case 2-D dims={dims}", flush=True)
    for npdc in ((3, 3, 3), (1, 2, 3), (4, 1, 3)):
        g, ref, lsh = run_both(dims, npdc, 3, True)
        assert np.array_equal(lsh, ref["lshock"])
        for m in range(5):
            e = nerr(g["qrhs"][..., m], ref["qrhs"][..., m])
            print(f"  npdc {npdc} qrhs{m + 1}: {e:.2e}")
            assert e <= 1e-13


@pytest.mark.parametrize("compact", [False, True])
def test_upwind_long_lines_two_dimensional(compact):
    """x lines longer than one 256-interface tile (several segments per line)
    and an axis beyond the compact line sweeps' 505-cell limit."""
    dims = (600, 40, 0)
    print(f"case long 2-D dims={dims} compact={compact}", flush=True)
    g, ref, lsh = run_both(dims, (1, 3, 3), 3, True, compact=compact)
    assert np.array_equal(lsh, ref["lshock"])
    for m in range(5):
        e = nerr(g["qrhs"][..., m], ref["qrhs"][..., m])
        print(f"  qrhs{m + 1}: {e:.2e}")
        assert e <= 1e-13


def test_upwind_with_compact_gradient():
    dims = (36, 34, 32)
    g, ref, lsh = run_both(dims, (3, 3, 3), 3, True, explicit=False)
    assert np.array_equal(lsh, ref["lshock"])
    for m in range(5):
        # the compact gradient enters only through the shock flags, which
        # agree; the convective fluxes themselves are the checker's
        assert nerr(g["qrhs"][..., m], ref["qrhs"][..., m]) <= 1e-13


@pytest.mark.parametrize("ntype", [1, 2, 3, 4])
def test_explicit_derivative_field(ntype):
    dims = (40, 36, 33)
    p = ag.default_params(dims, npdci=ntype, npdcj=ntype, npdck=ntype, deriv_explicit=1)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(200 + ntype)
    f = np.asfortranarray(rng.standard_normal(ag.haloed_shape(p)))
    for d in (1, 2, 3):
        df = s.central(d, f)
        ax = d - 1
        gm = np.moveaxis(f, ax, 0)
        ref = np.zeros_like(np.moveaxis(ag.interior(f, p), ax, 0))
        for a in range(HM, gm.shape[1] - HM):
            for b in range(HM, gm.shape[2] - HM):
                ref[:, a - HM, b - HM] = po.diff6ec(ntype, np.ascontiguousarray(gm[:, a, b]))
        ref = np.moveaxis(ref, 0, ax)
        assert np.array_equal(df, ref), (ntype, d, nerr(df, ref))
    s.close()


@pytest.mark.parametrize("convection", [1, 2])
def test_upwind_periodic_steps_conserve_and_match_pieces(convection):
    """3-D periodic Euler box stepped with an upwind scheme: the whole step
    equals the stage pieces; with the explicit scheme the unique-point sums
    of q (mass, momentum, energy) are conserved to rounding (the interface
    fluxes telescope; the compact scheme's non-cyclic periodic systems do
    not telescope exactly)."""
    dims = (32, 32, 32)
    p = ag.default_params(dims, convection=convection, deriv_explicit=1, recon_schem=3,
                          lchardecomp=1, lfilter=0, diffterm=0, mach=1.0)
    st = quadrant_state(dims, seed=3)
    # make the state exactly periodic: halos and the duplicate planes copied
    # from the unique points 0..n-1
    for k in st:
        a = st[k]
        for ax in range(3):
            n = dims[ax]
            idx = (np.arange(-HM, n + HM + 1) % n) + HM
            a = np.take(a, idx, axis=ax)
        st[k] = np.asfortranarray(a)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s1 = ag.AstrSolver(p)
    s2 = ag.AstrSolver(p)
    for s in (s1, s2):
        s.upload_geom(jac, dxi)
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])

    def totals(s):
        q = ag.interior(s.download_state()["q"], p)[:-1, :-1, :-1]
        return q.reshape(-1, 5).sum(axis=0)

    t0 = totals(s1)
    dt = 2e-4
    for k in range(3):
        s1.time_integration_rk(dt)
        for stage in (1, 2, 3):
            s2.qswap()
            s2.rhscal()
            s2.rk3_update(stage, dt)
    a, b = s1.download_state(), s2.download_state()
    for nm in a:
        assert np.array_equal(a[nm], b[nm]), nm
    t1 = totals(s1)
    scale = np.abs(ag.interior(a["q"], p)).reshape(-1, 5).sum(axis=0)
    print("conservation", (t1 - t0) / scale)
    if convection == 1:
        assert np.all(np.abs(t1 - t0) <= 1e-12 * scale)
    assert np.all(np.isfinite(t1))
    s1.close()
    s2.close()


def _lines(f, p, ax):
    """(a, b, line) over the interior lines of axis ax of a haloed field."""
    g = np.moveaxis(f, ax, 0)
    for a in range(HM, g.shape[1] - HM):
        for b in range(HM, g.shape[2] - HM):
            yield a, b, np.ascontiguousarray(g[:, a, b])


@pytest.mark.parametrize("ntype", [1, 2, 3, 4])
@pytest.mark.parametrize("scheme", ["compact", 0, 2, 5, 6])
def test_interface_flux_field_matches_checker(ntype, scheme):
    """astr_flux_field against the checker's flux_compact / recons_exp on
    every line of every axis, both winds: bit-equal."""
    dims = (24, 20, 18)
    compact = scheme == "compact"
    p = ag.default_params(dims, npdci=ntype, npdcj=ntype, npdck=ntype,
                          convection=2 if compact else 1, recon_schem=0 if compact else scheme,
                          lchardecomp=0, lfilter=0, diffterm=0, mach=1.0)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(300 + ntype)
    # a smooth part and a jump, so that the limiters act
    f = rng.standard_normal(ag.haloed_shape(p)) * 0.1
    f[ag.haloed_shape(p)[0] // 2:, :, :] += 1.0
    f = np.asfortranarray(f)
    for d in (1, 2, 3):
        ax = d - 1
        dim = dims[ax]
        for wind in (1, -1):
            shock = wind > 0
            fh = np.moveaxis(s.interface_flux(d, f, wind=wind, shock=shock), ax, 0)
            for a, b, line in _lines(f, p, ax):
                got = fh[HM - 1:HM + dim + 1, a, b]  # nodes -1..dim
                if compact:
                    ref = po.flux_compact(ntype, line, plus=wind > 0, bfac=p.bfacmpld)
                    assert np.array_equal(got, ref), (d, wind, a, b, nerr(got, ref))
                    continue
                lo = 0 if ntype in (1, 4) else -1
                hi = dim - 1 if ntype in (2, 4) else dim
                for i in range(lo, hi + 1):
                    n = i + HM
                    f8 = line[n - 3:n + 5] if wind > 0 else line[n - 3:n + 5][::-1]
                    nte = (1 if i < dim // 2 else 2) if ntype == 4 else ntype
                    ref = po.recons_exp(f8, i, dim, nte, scheme, shock, p.bfacmpld)
                    assert got[i + 1] == ref, (d, wind, a, b, i, got[i + 1], ref)
    s.close()


def test_interface_flux_field_rejects_central_context():
    p = ag.default_params((16, 16, 16))
    s = ag.AstrSolver(p)
    f = np.zeros(ag.haloed_shape(p), order="F")
    with pytest.raises(ag.AstrError):
        s.interface_flux(1, f)
    s.close()
