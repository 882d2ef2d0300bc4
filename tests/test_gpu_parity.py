"""Parity of the HIP solver (through the C ABI) with the checker.

Tolerances (written per test):
  * single derivative / filter sweeps: max |gpu - ref| <= 1e-12 * max |ref|
    over the field (the chunked-scan Thomas solve and FMA contraction change
    only the rounding);
  * stage pieces (filterq, qswap, rhscal): the same 1e-12 normwise bound;
    qrhs, a small difference of large pressure fluxes, within 1e-13 of the
    pressure-flux derivative scale max|p| * max d(xi)/dx, and the gradients
    within 1e-13 of max|f| * max d(xi)/dx of the differentiated field;
  * whole steps: KE and enstrophy within 1e-10 relative, primitive fields
    within 1e-9 absolute (the bar the reference sets for 100 RK3 steps).
"""
import math
import os

import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
HM = 5


def nerr(a, b):
    """max |a-b| / max |b| (normwise relative error of a field)."""
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def lines_ref(fn, f, direction):
    """Apply a 1-D line operator of the checker to every line of a haloed
    field along `direction` (1..3); f has shape (ni+11, nj+11, nk+11)."""
    ax = direction - 1
    g = np.moveaxis(f, ax, 0)
    other = g.shape[1:]
    out = np.zeros((g.shape[0] - 2 * HM,) + tuple(o - 2 * HM for o in other))
    for a in range(HM, other[0] - HM):
        for b in range(HM, other[1] - HM):
            out[:, a - HM, b - HM] = fn(np.ascontiguousarray(g[:, a, b]))
    return np.moveaxis(out, 0, ax)


@pytest.mark.parametrize("ntype", [1, 2, 3, 4])
def test_field_line_operators(ntype):
    dims = (40, 36, 33)
    p = ag.default_params(dims, npdci=ntype, npdcj=ntype, npdck=ntype)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(100 + ntype)
    f = np.asfortranarray(rng.standard_normal(ag.haloed_shape(p)))
    for d in (1, 2, 3):
        df = s.central(d, f)
        ref = lines_ref(lambda ln: pyoracle.df_compact(ntype, ln), f, d)
        e = nerr(df, ref)
        print(f"deriv ntype {ntype} dir {d}: {e:.2e}")
        assert e < 1e-12
        ff = s.compact_filter(d, f)
        ref = lines_ref(lambda ln: pyoracle.compact_filter(ntype, ln), f, d)
        e = nerr(ff, ref)
        print(f"filter ntype {ntype} dir {d}: {e:.2e}")
        assert e < 1e-12
    s.close()


def test_field_line_operators_long_lines():
    # 256-node lines (5 rows per lane) against the golden vectors of the
    # reference's own modules
    G = np.load(os.path.join(GOLDEN, "lines_ref.npz"))
    p = ag.default_params((256, 16, 16))
    s = ag.AstrSolver(p)
    f = np.zeros(ag.haloed_shape(p), order="F")
    f[:, :, :] = G["f_smooth_3_256"][:, None, None]
    df = s.central(1, f)
    ff = s.compact_filter(1, f)
    assert nerr(df[:, 3, 4], G["df_smooth_3_256"]) < 1e-12
    assert nerr(ff[:, 3, 4], G["ff098_smooth_3_256"]) < 1e-12
    s.close()


@pytest.mark.parametrize("ntype", [3, 4])
def test_field_line_operators_longest_lines(ntype):
    """505-cell y and z lines (16 rows per lane): the y/z tiles take 69 KB of
    LDS, beyond the 64 KB of earlier GPUs; against the checker line by line."""
    dims = (16, 505, 505) if ntype == 3 else (16, 480, 16)
    p = ag.default_params(dims, npdci=3, npdcj=ntype, npdck=ntype)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(7 + ntype)
    f = np.asfortranarray(rng.standard_normal(ag.haloed_shape(p)))
    for d in ((2, 3) if ntype == 3 else (2,)):
        df = s.central(d, f)
        ref = lines_ref(lambda ln: pyoracle.df_compact(ntype, ln), f, d)
        e = nerr(df, ref)
        print(f"deriv ntype {ntype} dir {d} n {dims[d - 1]}: {e:.2e}", flush=True)
        assert e < 1e-12
        ff = s.compact_filter(d, f)
        ref = lines_ref(lambda ln: pyoracle.compact_filter(ntype, ln), f, d)
        e = nerr(ff, ref)
        print(f"filter ntype {ntype} dir {d} n {dims[d - 1]}: {e:.2e}", flush=True)
        assert e < 1e-12
    s.close()


def test_axis_beyond_the_line_operators_is_rejected():
    with pytest.raises(ag.AstrError):
        ag.AstrSolver(ag.default_params((16, 506, 16)))


@pytest.mark.parametrize("npdc", [(4, 4, 4), (4, 3, 4), (3, 3, 4)])
def test_bounded_axes_central_steps_match_checker(npdc):
    """Whole RK3 steps of the central path with bounded axes (the compact
    closures of npdc 4 at both faces of the axis, qswap only on the periodic
    axes) against the checker: statistics at 1e-10, every node and the
    periodic axes' halo faces at 1e-10 after 3 steps."""
    dims = (40, 36, 34)
    print(f"case bounded npdc={npdc} dims={dims}", flush=True)
    ob = pyoracle.OracleBox(dims, variant=0, npdc=npdc)
    ob.init_tgv()
    st = ob.state_haloed()
    p = ag.default_params(dims, npdci=npdc[0], npdcj=npdc[1], npdck=npdc[2])
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    dt = 0.004
    for n in range(3):
        so = ob.step(dt, stats=True)
        sg = s.time_integration_rk(dt, stats=True)
        for a, b in zip(sg, so):
            assert abs(a - b) <= 1e-10 * abs(b), (n, sg, so)
    print(f"  3 steps {compare_state(ob, s, p, 1e-10, halo_faces=True, vector_scale=True):.2e}")
    # the boundary nodes moved: a bounded face is no periodic seam
    q = ag.interior(s.download_state()["q"][..., 1], p)
    if npdc[0] != 3:
        assert np.max(np.abs(q[0] - q[-1])) > 1e-6
    s.close()


@pytest.mark.parametrize("v", [0.2, -1.2])
def test_gcnscbc_central_steps_match_checker(v):
    """gcnscbc (bctype 23) on the x-high and y-high faces of a three-dimensional
    box (bounded x and y, periodic z) through whole RK3 steps of the central
    path against the checker: statistics at 1e-10, every node and the
    periodic axis' halo faces at 1e-10 after 3 steps.  Ma 0.4 keeps the
    integer sound speed gcnscbc takes (sqrt(T)/Ma near 2.5) away from a
    truncation boundary; v = -1.2 takes face 4's inflow branch (integer
    normal velocity -1)."""
    from test_oracle_bounded import GCN_FACES, pulse

    dims, mach = (32, 28, 24), 0.4
    print(f"case gcnscbc central dims={dims} v={v}", flush=True)
    ob = pyoracle.OracleBox(dims, variant=0, npdc=(4, 4, 3), mach=mach)
    ob.set_bc(GCN_FACES)
    pulse(ob, mach, v=v)
    st = ob.state_haloed()
    p = ag.default_params(dims, npdci=4, npdcj=4, npdck=3, mach=mach)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.set_lengths(2 * np.pi, 2 * np.pi, 2 * np.pi)
    s.set_bc(tuple(23 if f == 23 else 1 for f in GCN_FACES), 0)
    dt = 0.01
    for n in range(3):
        so = ob.step(dt, stats=True)
        sg = s.time_integration_rk(dt, stats=True)
        for a, b in zip(sg, so):
            assert abs(a - b) <= 1e-10 * abs(b), (n, sg, so)
    print(f"  3 steps {compare_state(ob, s, p, 1e-10, halo_faces=True, vector_scale=True):.2e}")
    s.close()


@pytest.mark.parametrize("faces", [(50, 50, 50, 50, 0, 0), (60, 60, 60, 60, 0, 0),
                                   (60, 50, 50, 60, 0, 0), (41, 41, 42, 42, 0, 0),
                                   (50, 21, 50, 21, 0, 0), (50, 50, 31, 31, 0, 0),
                                   (50, 21, 421, 421, 0, 0), (50, 50, 411, 411, 0, 0)])
def test_zeroextrap_central_steps_match_checker(faces):
    """zeroextrap (bctype 50), symmetry (60) and the noslip walls (41, 42) on
    the four faces of the bounded x and y axes of a three-dimensional box
    (periodic z; the walls make it a viscous channel) through
    whole RK3 steps of the central path against the checker, applied face by
    face before each stage's qswap: statistics at 1e-10, every node and the
    periodic axis' halo faces at 1e-10 after 3 steps."""
    from test_oracle_bounded import pulse

    dims, mach = (32, 28, 24), 0.4
    print(f"case zeroextrap/symmetry central dims={dims} faces={faces}", flush=True)
    tw = (1.05, 0.95, 1.1, 0.9, 1.0, 1.0)
    ob = pyoracle.OracleBox(dims, variant=0, npdc=(4, 4, 3), mach=mach)
    ob.set_bc(faces)
    ob.set_twall(tw)
    ob.set_xslip(math.pi)  # x(16) = 2 pi/32*16 is pi exactly: it slips
    pulse(ob, mach)
    st = ob.state_haloed()
    p = ag.default_params(dims, npdci=4, npdcj=4, npdck=3, mach=mach)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.set_twall(tw)
    s.set_lengths(2 * math.pi, 2 * math.pi, 2 * math.pi)  # the checker box's
    s.set_xslip(math.pi)
    s.set_bc(tuple(f if f else 1 for f in faces), 0)
    dt = 0.01
    for n in range(3):
        so = ob.step(dt, stats=True)
        sg = s.time_integration_rk(dt, stats=True)
        for a, b in zip(sg, so):
            assert abs(a - b) <= 1e-10 * abs(b), (n, sg, so)
    print(f"  3 steps {compare_state(ob, s, p, 1e-10, halo_faces=True, vector_scale=True):.2e}")
    s.close()


def test_gcnscbc_rejected_by_batch_and_adjoint():
    dims = (16, 16, 16)
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.tgv_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.set_bc((1, 23, 1, 1, 1, 1), 0)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        s.adjoint_enstrophy(2, 0.001, 2)
    with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
        ag.AstrBatch([s])
    s.close()


def test_bounded_axes_rejected_where_not_built():
    dims = (32, 32, 32)
    for over in (dict(npdci=4, convection=1, diffterm=0), dict(npdcj=4, variant=1),
                 dict(npdck=1), dict(npdci=2)):
        p = ag.default_params(dims, **over)
        s = ag.AstrSolver(p)
        s.upload_geom(*ag.box_metrics(p))
        if p.variant == 1:
            s.set_spacing(2 * np.pi / dims[0], 2 * np.pi / dims[1], 2 * np.pi / dims[2])
        st = ag.tgv_state(ag.default_params(dims))
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        with pytest.raises(ag.AstrError, match="ASTR_E_UNSUPPORTED"):
            s.time_integration_rk(0.001)
        s.close()


def make_pair(dims):
    ob = pyoracle.OracleBox(dims, variant=0)
    ob.init_tgv()
    st = ob.state_haloed()
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    return ob, s, p


PRIMS = [("rho", None), ("u", 0), ("v", 1), ("w", 2), ("prs", None), ("tmp", None)]


def compare_state(ob, s, p, tol, halo_faces=False, vector_scale=False):
    """Every node of q and the primitives against the checker, max |d| over
    the field's max |ref|; vector_scale: the momentum components against the
    momentum vector's max, the velocity components against the velocity's
    (a component that starts at zero, as rho w of the Taylor-Green field, is
    O(dt) after one step, and its own max is no scale for rounding)."""
    out = s.download_state()
    worst = 0.0
    ref_q = [ob.get(f"q{m + 1}") for m in range(5)]
    ref_p = {nm: ob.get(nm) for nm, _ in PRIMS}
    mom = max(float(np.max(np.abs(ref_q[m]))) for m in (1, 2, 3))
    vel = max(float(np.max(np.abs(ref_p[nm]))) for nm in "uvw")

    def err(a, b, vscale):
        if vector_scale and vscale:
            return float(np.max(np.abs(a - b))) / vscale
        return nerr(a, b)

    for m in range(5):
        e = err(ag.interior(out["q"][..., m], p), ref_q[m], mom if m in (1, 2, 3) else None)
        worst = max(worst, e)
        assert e < tol, (f"q{m + 1}", e)
    for nm, comp in PRIMS:
        g = out["vel"][..., comp] if comp is not None else out[nm]
        e = err(ag.interior(g, p), ref_p[nm], vel if nm in "uvw" else None)
        worst = max(worst, e)
        assert e < tol, (nm, e)
    if halo_faces:
        # the halo cells qswap writes: i-halos over j,k in 0..n etc.
        n0, n1, n2 = p.im, p.jm, p.km
        I0, I1, I2 = slice(HM, HM + n0 + 1), slice(HM, HM + n1 + 1), slice(HM, HM + n2 + 1)
        lo, hi = slice(0, HM), slice(-HM, None)
        faces = [(lo, I1, I2), (hi, I1, I2), (I0, lo, I2), (I0, hi, I2), (I0, I1, lo),
                 (I0, I1, hi)]
        for m in range(5):
            ref = ob.get_haloed(f"q{m + 1}")
            for fc in faces:
                e = err(out["q"][..., m][fc], ref[fc], mom if m in (1, 2, 3) else None)
                assert e < tol, ("q halo", m, e)
        for nm, comp in PRIMS:
            ref = ob.get_haloed(nm)
            g = out["vel"][..., comp] if comp is not None else out[nm]
            for fc in faces:
                e = err(g[fc], ref[fc], vel if nm in "uvw" else None)
                assert e < tol, (nm + " halo", e)
    return worst


def test_stage_pieces_match_oracle():
    dims = (40, 36, 32)
    ob, s, p = make_pair(dims)
    ob.filterq()
    s.filterq()
    print("filterq", compare_state(ob, s, p, 1e-12))
    ob.qswap()
    s.qswap()
    print("qswap", compare_state(ob, s, p, 1e-12, halo_faces=True))
    ob.rhscal()
    s.rhscal()
    r = s.download_rhs()
    # The residual is a small difference of fluxes of size p * d(xi)/dx
    # (p ~ 1/(gamma Ma^2) = 71, the Taylor-Green residual is ~1e-3 and its
    # mass component ~1e-10), so flux rounding sets an absolute floor: the
    # bound is 1e-13 of the pressure-flux derivative scale.
    flux_scale = float(np.max(np.abs(ob.get("prs")))) * max(p.im, p.jm, p.km) / (2 * np.pi)
    for m in range(5):
        e = float(np.max(np.abs(r["qrhs"][..., m] - ob.deriv(12 + m)))) / flux_scale
        print(f"qrhs{m + 1}: {e:.2e}")
        assert e < 1e-13
    # gradients against the derivative scale of the differentiated field,
    # max|f| * max d(xi)/dx (T ~ 1 varies by 1e-3 in this field)
    xi = max(p.im, p.jm, p.km) / (2 * np.pi)
    fs = max(float(np.max(np.abs(ob.get(nm)))) for nm in "uvw") * xi
    for m in range(3):
        for d in range(3):
            e = float(np.max(np.abs(r["dvel"][..., m, d] - ob.deriv(3 * m + d)))) / fs
            assert e < 1e-13, ("dvel", m, d, e)
    fs = float(np.max(np.abs(ob.get("tmp")))) * xi
    for d in range(3):
        e = float(np.max(np.abs(r["dtmp"][..., d] - ob.deriv(9 + d)))) / fs
        assert e < 1e-13, ("dtmp", d, e)
    st_o = ob.stats()
    st_g = s.statcal()
    for a, b in zip(st_g, st_o):
        assert abs(a - b) <= 1e-12 * abs(b)
    s.close()
    ob.close()


def test_steps_match_oracle():
    dims = (32, 32, 32)
    ob, s, p = make_pair(dims)
    for k in range(5):
        so = ob.step(0.004, stats=True)
        sg = s.time_integration_rk(0.004, stats=True)
        for a, b in zip(sg, so):
            assert abs(a - b) <= 1e-10 * abs(b), (k, sg, so)
    print("5 steps", compare_state(ob, s, p, 1e-10, halo_faces=True))
    s.close()
    ob.close()


@pytest.mark.parametrize("dims", [(40, 36, 33), (45, 43, 27)])
def test_filterq_keeps_every_cell_outside_the_nodes(dims):
    """filterq writes the nodes 0..n only; the sweeps also store the rest of
    the 128-byte line that holds node n (so that HBM sees whole lines), with
    the values already there.  Every cell outside 0..n along some axis (halo
    faces, edges and corners) must come back bit for bit."""
    print(f"case filterq keeps dims={dims}", flush=True)
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    rng = np.random.default_rng(4)
    shp = ag.haloed_shape(p)
    q = np.asfortranarray(1.0 + 0.1 * rng.standard_normal(shp + (5,)))
    rho = np.asfortranarray(q[..., 0].copy())
    vel = np.asfortranarray(0.1 * rng.standard_normal(shp + (3,)))
    prs = np.asfortranarray(1.0 + 0.1 * rng.standard_normal(shp))
    tmp = np.asfortranarray(1.0 + 0.1 * rng.standard_normal(shp))
    s.upload_state(q, rho, vel, prs, tmp)
    s.filterq()
    out = s.download_state()
    inside = np.zeros(shp, dtype=bool)
    inside[HM:HM + dims[0] + 1, HM:HM + dims[1] + 1, HM:HM + dims[2] + 1] = True
    for m in range(5):
        assert np.array_equal(out["q"][..., m][~inside], q[..., m][~inside]), m
        assert not np.array_equal(out["q"][..., m][inside], q[..., m][inside]), m
    assert np.array_equal(out["vel"], vel) and np.array_equal(out["prs"], prs)
    s.close()


def test_stage_pieces_reproduce_the_fused_step():
    """The whole step applies the RK3 combine inside the last flux sweep;
    driving the same stage through the separate entry points (filterq,
    qswap, rhscal, rk3_update) does the same arithmetic in the same order.
    The compiler may contract a multiply-add differently in the two
    kernels, so the bound is 1e-14 of each field's magnitude rather than
    bit equality (a wrong combine or a missed update is off by far more)."""
    dims = (36, 34, 32)
    _, s1, p = make_pair(dims)
    _, s2, _ = make_pair(dims)
    for k in range(2):
        st1 = s1.time_integration_rk(0.004, stats=True)
        for stage in (1, 2, 3):
            s2.filterq()
            s2.qswap()
            s2.rhscal()
            if stage == 1:
                st2 = s2.statcal()
            s2.rk3_update(stage, 0.004)
        for x, y in zip(st1, st2):
            assert abs(x - y) <= 1e-14 * abs(y), (k, st1, st2)
        a, b = s1.download_state(), s2.download_state()
        for nm in a:
            e = nerr(a[nm], b[nm])
            print(f"step {k} {nm}: {'bit-equal' if np.array_equal(a[nm], b[nm]) else e}")
            assert e <= 1e-14, (k, nm, e)
    s1.close()
    s2.close()


@pytest.mark.parametrize("dims", [(48, 40, 36), (45, 43, 29)])
def test_non_cubic_steps_match_oracle(dims):
    """(45, 43, 29): the x halo strip n+1..n+5 runs past the 128-byte line
    that holds node n (n mod 16 = 13), and the end rows of the flux
    derivative read G_x up to n+4, beyond that line."""
    print(f"case non-cubic steps dims={dims}", flush=True)
    ob, s, p = make_pair(dims)
    for k in range(3):
        ob.step(0.004)
    s.steploop(3, 0.004)
    print("non-cubic 3 steps", compare_state(ob, s, p, 1e-10))
    s.close()
    ob.close()


def test_tgv64_100_steps_against_checker_series():
    """64^3 Taylor-Green box, 100 RK3 steps of dt 0.004 (the small-size
    workload): KE and enstrophy within 1e-10 relative at every step,
    primitive fields within 1e-9 absolute at the probes."""
    G = np.load(os.path.join(GOLDEN, "tgv64_main.npz"))
    p = ag.default_params(64)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    ob = pyoracle.OracleBox(64, variant=0)
    ob.init_tgv()
    st = ob.state_haloed()
    ob.close()
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    stats = s.steploop(100, 0.004, stats=True)
    ref = G["stats"]
    ke_err = np.max(np.abs(stats[:, 0] - ref[:, 0]) / np.abs(ref[:, 0]))
    en_err = np.max(np.abs(stats[:, 1] - ref[:, 1]) / np.abs(ref[:, 1]))
    print(f"KE rel err {ke_err:.2e}, enstrophy rel err {en_err:.2e}")
    assert ke_err < 1e-10 and en_err < 1e-10
    out = s.download_state()
    for nm, comp in PRIMS:
        g = out["vel"][..., comp] if comp is not None else out[nm]
        g = ag.interior(g, p)[::8, ::8, ::8]
        e = float(np.max(np.abs(g - G["probe_" + nm])))
        print(f"{nm} probe abs err {e:.2e}")
        assert e < 1e-9
    s.close()


def test_curvilinear_metrics_rejected_where_not_built():
    for over in (dict(convection=1, diffterm=0), dict(variant=1)):
        p = ag.default_params(32, **over)
        s = ag.AstrSolver(p)
        jac, dxi = ag.box_metrics(p)
        dxi[10, 10, 10, 0, 1] = 1e-3
        with pytest.raises(ag.AstrError, match="UNSUPPORTED"):
            s.upload_geom(jac, dxi)
        s.close()


def test_fortran_host_drives_the_solver():
    """The iso_c_binding module (astr_gpu_amd/fortran/astr_hip_mod.f90) used
    from a Fortran program gives the same statistics as the Python binding
    on the same Taylor-Green box (initial fields from each host's libm)."""
    import subprocess
    from astr_gpu_amd.build import FORTRAN_EXE
    if not os.path.exists(FORTRAN_EXE):
        pytest.skip("Fortran driver not built (no amdflang where build() ran)")
    out = subprocess.run([FORTRAN_EXE, "32", "3", "0.004"], capture_output=True, text=True,
                         timeout=60, check=True).stdout.split("\n")
    f_stats = np.array([[float(v) for v in ln.split()[1:]] for ln in out if ln.strip()
                        and not ln.strip().startswith("max")])
    p = ag.default_params(32)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.tgv_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    py_stats = s.steploop(3, 0.004, stats=True)
    s.close()
    assert f_stats.shape == (3, 3)
    assert np.max(np.abs(f_stats - py_stats) / np.abs(py_stats)) < 1e-12


def mini_pair(n):
    ob = pyoracle.OracleBox(n, variant=1)
    ob.init_tgv()
    st = ob.state_haloed()
    p = ag.default_params(n, variant=1, sutherland=110.4, beta_bound=1.09)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    h = 2 * np.pi / n
    s.set_spacing(h, h, h)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    return ob, s, p


def test_miniapp_variant_steps_match_oracle():
    """Mini-app stage order (tgvsolver.F90: bchomo, df/dx, filter after the
    update) on 32^3 against the checker's variant, whole fields."""
    ob, s, p = mini_pair(32)
    for k in range(4):
        so = ob.step(1e-3, stats=True)
        sg = s.time_integration_rk(1e-3, stats=True)
        for a, b in zip(sg[:2], so[:2]):
            assert abs(a - b) <= 1e-12 * abs(b), (k, sg, so)
    print("mini 4 steps", compare_state(ob, s, p, 1e-10))
    s.close()
    ob.close()


def test_miniapp_variant_reproduces_reference_state_ref_128():
    """The reference-held series miniapps/tgv_solver_3d/state.ref_128: KE and
    enstrophy of the reference's own mini-app at every one of 100 steps
    (128^3, dt 1e-3).  The GPU mini-app variant stays within 1e-10 relative
    of the reference's numbers (its summation order and the chunked-scan
    solves differ only in rounding)."""
    ref = np.loadtxt(os.path.join(GOLDEN, "state.ref_128.txt"))
    ob, s, p = mini_pair(128)
    ob.close()
    st = np.asarray(s.steploop(ref.shape[0], 1e-3, stats=True))
    eke = np.max(np.abs(st[:, 0] - ref[:, 2]) / ref[:, 2])
    een = np.max(np.abs(st[:, 1] - ref[:, 3]) / ref[:, 3])
    print(f"state.ref_128: KE max rel {eke:.2e}, enstrophy max rel {een:.2e}")
    assert eke <= 1e-10 and een <= 1e-10
    s.close()
