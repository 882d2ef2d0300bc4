"""Pins of the checker's bounded-box upwind pieces (oracle/upwind.c): the
ntype-4 closure rule, the 2-D Riemann example's boundary condition, the
single-rank qswap and the upwind RK3 step.

No reference test stores outputs of these routines and the reference's
full program needs MPI, so each is pinned by what the mathematics or the
reference's own code paths fix:
  * ntype 4 (one domain bounded at both ends) equals ntype 1 on its low
    half and ntype 2 on its high half, node for node, and a two-rank split
    of the axis (ntype 1 | ntype 2 with halos from the neighbour, the only
    decomposition on which the reference runs a bounded axis) reproduces
    the one-domain right-hand side bit for bit;
  * shockspeed is the Rankine-Hugoniot shock speed: the normal-shock
    pressure ratio of its Mach number returns p2/p1;
  * the boundary states at time 0 are the initial quadrant states of
    riem2dini on faces 1 and 3; the extrapolated faces set the one-sided
    second-order slope to zero (exact for a quadratic with zero slope);
  * qswap of a 2-D box: every k plane equals the plane k = 0 and the
    primitives equal q2fvar(q); periodic axes: halos are the periodic
    images and the duplicate nodes their average;
  * an RK3 step keeps a uniform state uniform.
"""
import numpy as np
import pytest

import pyoracle as po

HM = 5
GAMMA = 1.4


def bounded_quadrants(dims, mach=1.0, seed=0):
    """Four-quadrant states split at 0.8 with a smooth perturbation; halos
    filled with the same closed form (bounded axes never read them)."""
    im, jm, km = dims
    rng = np.random.default_rng(seed)
    t = [np.arange(-HM, n + HM + 1) / max(n, 1) for n in dims]
    X, Y, Z = np.meshgrid(*t, indexing="ij")
    ne, nw = (X >= 0.8) & (Y >= 0.8), (X < 0.8) & (Y >= 0.8)
    sw, se = (X < 0.8) & (Y < 0.8), (X >= 0.8) & (Y < 0.8)
    rho = np.select([ne, nw, sw, se], [1.5, 0.5323, 0.138, 0.5323])
    u = np.select([ne, nw, sw, se], [0.0, 1.206, 1.206, 0.0])
    v = np.select([ne, nw, sw, se], [0.0, 0.0, 1.206, 1.206])
    p = np.select([ne, nw, sw, se], [1.5, 0.3, 0.029, 0.3])
    pert = 0.03 * np.sin(2 * np.pi * (X + 2 * Y) + rng.random())
    return po.prim_to_state(rho * (1 + pert), u + 0.1 * pert, v - 0.05 * pert, 0 * X,
                            p * (1 + 0.5 * pert), mach=mach)


def metrics(dims):
    hs = tuple(n + 1 + 2 * HM for n in dims)
    dxi = np.zeros(hs + (9,), order="F")
    h = []
    for d, n in enumerate(dims):
        h.append(1.0 / n if n else 1.0)
        dxi[..., 4 * d] = 1.0 / h[-1]
    return dxi, np.full(hs, h[0] * h[1] * h[2], order="F")


def rhs(dims, npdc, st, lchar=True, recon=3, geom=None):
    box = po.UwdBox(dims, npdc=npdc, recon_schem=recon, lchardecomp=lchar, shkcrt=0.01)
    dxi, jac = geom if geom is not None else metrics(dims)
    return box.rhscal(st, dxi, jac)


@pytest.mark.parametrize("recon", [3, 5, 0])
def test_ntype4_is_ntype1_below_and_ntype2_above(recon):
    dims = (48, 20, 0)
    st = bounded_quadrants(dims)
    r4 = rhs(dims, (4, 3, 3), st, recon=recon)
    r1 = rhs(dims, (1, 3, 3), st, recon=recon)
    r2 = rhs(dims, (2, 3, 3), st, recon=recon)
    n = dims[0]
    # the three boxes differ only at the far end of each one-sided type
    # (closure rows, clamps and the -4..+5 sensor window reach 16 nodes)
    assert np.array_equal(r4["qrhs"][:n - 16], r1["qrhs"][:n - 16])
    assert np.array_equal(r4["qrhs"][16:], r2["qrhs"][16:])
    assert np.array_equal(r4["lshock"][:n - 16], r1["lshock"][:n - 16])
    assert np.array_equal(r4["lshock"][16:], r2["lshock"][16:])
    # and they do differ there: the ends are treated as bounded
    assert not np.array_equal(r4["qrhs"], r1["qrhs"])
    assert not np.array_equal(r4["qrhs"], r2["qrhs"])


@pytest.mark.parametrize("axis", [0, 1])
def test_ntype4_equals_a_two_rank_split(axis):
    """The reference runs a bounded axis on several ranks only: ntype 1 on
    the first, 2 on the last, halos from the neighbour.  The explicit upwind
    flux has no closure at a rank interface, so the split reproduces the
    one-domain ntype-4 result node for node (sensor off: its ssf halo swap
    between ranks is not modelled here)."""
    dims = (48, 40, 0)
    st = bounded_quadrants(dims, seed=2)
    npdc = [3, 3, 3]
    npdc[axis] = 4
    whole = rhs(dims, tuple(npdc), st, lchar=False)["qrhs"]
    n = dims[axis]
    cut = n // 2
    dxi, jac = metrics(dims)
    parts = []
    for lo, hi, nt in ((0, cut, 1), (cut, n, 2)):
        def cut_out(a):
            idx = [slice(None)] * a.ndim
            idx[axis] = slice(lo, hi + 2 * HM + 1)
            return np.asfortranarray(a[tuple(idx)])
        sub = {k: cut_out(a) for k, a in st.items()}
        sd = list(dims)
        sd[axis] = hi - lo
        npd = [3, 3, 3]
        npd[axis] = nt
        parts.append(rhs(tuple(sd), tuple(npd), sub, lchar=False,
                         geom=(cut_out(dxi), cut_out(jac)))["qrhs"])
    lo_part = np.take(whole, range(0, cut + 1), axis=axis)
    hi_part = np.take(whole, range(cut, n + 1), axis=axis)
    assert np.array_equal(lo_part, parts[0])
    assert np.array_equal(hi_part, parts[1])


def test_shockspeed_is_the_rankine_hugoniot_speed():
    lib = po.uwd_lib()
    for mach in (1.0, 10.0):
        for p1, r1, p2 in ((0.3, 0.5323, 0.029), (1.0, 1.0, 4.5), (0.029, 0.138, 0.3)):
            t1 = p1 / r1 * GAMMA * mach * mach
            vs = lib.orc_riem2d_shockspeed(GAMMA, mach, p1, t1, p2)
            c1 = np.sqrt(GAMMA * p1 / r1)  # the sound speed in physical form
            ms = vs / c1
            # normal-shock pressure ratio of shock Mach number ms
            ratio = (2 * GAMMA * ms ** 2 - (GAMMA - 1)) / (GAMMA + 1)
            assert abs(ratio - p2 / p1) < 1e-13 * max(1.0, p2 / p1)


def test_boundary_states_at_time_zero_are_the_initial_quadrants():
    import astr_gpu_amd as ag
    dims = (64, 56, 0)
    p = ag.default_params(dims, mach=1.0)
    st = ag.riemann2d_state(p)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=p.gamma, mach=p.mach)
    o = box.riem2d_bc(st, 0.0)
    c = (slice(None), slice(None), HM)
    for k in ("rho", "prs", "tmp"):
        a, b = o[k][c], st[k][c]
        assert np.array_equal(a[HM, HM:-HM], b[HM, HM:-HM]), k       # face 1 (i = 0)
        assert np.array_equal(a[HM:-HM, HM], b[HM:-HM, HM]), k       # face 3 (j = 0)
    assert np.array_equal(o["vel"][HM, HM:-HM, HM], st["vel"][HM, HM:-HM, HM])
    assert np.array_equal(o["vel"][HM:-HM, HM, HM], st["vel"][HM:-HM, HM, HM])
    im, jm = dims[0] + HM, dims[1] + HM
    # the corner (im, jm) holds the north-east state at rest
    assert o["rho"][im, jm, HM] == 1.5 and o["prs"][im, jm, HM] == 1.5
    assert np.all(o["vel"][im, jm, HM] == 0.0)
    # nothing but the four faces changed
    inner = (slice(HM + 1, im), slice(HM + 1, jm), HM)
    for k in st:
        assert np.array_equal(o[k][inner], st[k][inner])


def test_extrapolated_faces_have_zero_one_sided_slope():
    dims = (40, 36, 0)
    im, jm = dims[0] + HM, dims[1] + HM
    t = np.arange(-HM, dims[0] + HM + 1, dtype=float)
    s = np.arange(-HM, dims[1] + HM + 1, dtype=float)
    T, S, _ = np.meshgrid(t, s, np.zeros(2 * HM + 1), indexing="ij")
    # quadratics with zero slope at i = im (for face 2) and j = jm (face 4)
    rho = (1.2 + 0.01 * (T - dims[0]) ** 2 * 2.0 ** -4 + 0.003 * (S - dims[1]) ** 2 * 2.0 ** -4)
    prs = 0.9 + 0.002 * (T - dims[0]) ** 2 + 0.001 * (S - dims[1]) ** 2
    u = 0.25 + 0.004 * (S - dims[1]) ** 2
    v = -0.3 + 0.002 * (T - dims[0]) ** 2
    st = po.prim_to_state(rho, u, v, 0 * rho, prs)
    box = po.UwdBox(dims, npdc=(4, 4, 3))
    o = box.riem2d_bc(st, 0.37, bctype=(1, 0, 1, 1, 1, 1))
    # face 2: rho, p, v extrapolated along i (exactly, the data being
    # quadratic with zero slope at im), u = 0
    for k, ref in (("rho", rho), ("prs", prs)):
        got = o[k][im, HM:jm, HM]
        assert np.allclose(got, ref[im, HM:jm, HM], rtol=0, atol=1e-15), k
    assert np.allclose(o["vel"][im, HM:jm, HM, 1], v[im, HM:jm, HM], rtol=0, atol=1e-15)
    assert np.all(o["vel"][im, HM:jm, HM, 0] == 0.0)
    o = box.riem2d_bc(st, 0.37, bctype=(1, 1, 1, 0, 1, 1))
    for k, ref in (("rho", rho), ("prs", prs)):
        assert np.allclose(o[k][HM:im, jm, HM], ref[HM:im, jm, HM], rtol=0, atol=1e-15), k
    assert np.allclose(o["vel"][HM:im, jm, HM, 0], u[HM:im, jm, HM], rtol=0, atol=1e-15)
    assert np.all(o["vel"][HM:im, jm, HM, 1] == 0.0)
    # a linear profile is not kept: the condition is a zero slope, not an
    # extrapolation of the slope
    lin = po.prim_to_state(1.0 + 0.01 * T, 0 * T, 0 * T, 0 * T, 1.0 + 0 * T)
    o = box.riem2d_bc(lin, 0.0, bctype=(1, 0, 1, 1, 1, 1))
    assert abs(o["rho"][im, HM + 3, HM] - (1.0 + 0.01 * (dims[0] - 2.0 / 3.0))) < 1e-14


def test_the_boundary_states_move_with_the_shock_speed():
    dims = (64, 64, 0)
    import astr_gpu_amd as ag
    p = ag.default_params(dims, mach=1.0)
    st = ag.riemann2d_state(p)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=p.gamma, mach=p.mach)
    lib = po.uwd_lib()
    c2 = p.gamma * p.mach ** 2
    p01 = float(np.float32(0.30))
    vss = lib.orc_riem2d_shockspeed(p.gamma, p.mach, p01, p01 / 0.5323 * c2, 0.029)
    time = 0.1
    o = box.riem2d_bc(st, time)
    y = np.arange(dims[1] + 1) / dims[1]
    upper = y > 0.8 - vss * time
    assert np.array_equal(o["rho"][HM, HM:HM + dims[1] + 1, HM] == 0.5323, upper)
    assert np.array_equal(o["rho"][HM, HM:HM + dims[1] + 1, HM] == 0.138, ~upper)


def test_qswap_of_a_two_dimensional_box():
    dims = (24, 20, 0)
    st = bounded_quadrants(dims, seed=5)
    # make the k planes differ so the copy is visible
    for k in st:
        st[k] = st[k] * (1.0 + 0.01 * np.arange(2 * HM + 1).reshape(
            (1, 1, -1) + (1,) * (st[k].ndim - 3)))
    box = po.UwdBox(dims, npdc=(4, 3, 3))
    o = box.qswap(st)
    jm = dims[1]
    inner_i = slice(HM, HM + dims[0] + 1)
    q = o["q"][inner_i]
    for kk in range(2 * HM + 1):
        assert np.array_equal(q[:, HM:HM + jm + 1, kk], q[:, HM:HM + jm + 1, HM])
    # j periodic: halos are the images, the duplicate rows their average
    qs = st["q"]
    avg = 0.5 * (qs[inner_i, HM, HM] + qs[inner_i, HM + jm, HM])
    assert np.array_equal(q[:, HM, HM], avg) and np.array_equal(q[:, HM + jm, HM], avg)
    assert np.array_equal(q[:, HM - 1, HM], qs[inner_i, HM + jm - 1, HM])
    assert np.array_equal(q[:, HM + jm + 1, HM], qs[inner_i, HM + 1, HM])
    # primitives are q2fvar(q) on the plane
    rho = q[:, HM:HM + jm + 1, HM, 0]
    u = q[:, HM:HM + jm + 1, HM, 1] / rho
    assert np.array_equal(o["rho"][inner_i, HM:HM + jm + 1, HM], rho)
    assert np.array_equal(o["vel"][inner_i, HM:HM + jm + 1, HM, 0], u)
    # the bounded i axis keeps its halos
    assert np.array_equal(o["q"][:HM], st["q"][:HM])


def test_rk3_step_keeps_a_uniform_state():
    dims = (32, 24, 0)
    one = np.ones(tuple(n + 1 + 2 * HM for n in dims))
    st = po.prim_to_state(1.1 * one, 0.4 * one, -0.3 * one, 0 * one, 0.8 * one)
    box = po.UwdBox(dims, npdc=(4, 4, 3), recon_schem=3, lchardecomp=True)
    dxi, jac = metrics(dims)
    o = box.rk3_steps(st, dxi, jac, 1e-3, nsteps=2)
    for k in st:
        assert np.allclose(o[k], st[k], rtol=1e-14, atol=0), k


def test_rk3_steps_of_the_riemann_problem_stay_physical():
    import astr_gpu_amd as ag
    dims = (48, 48, 0)
    p = ag.default_params(dims, mach=1.0)
    st = ag.riemann2d_state(p)
    box = po.UwdBox(dims, npdc=(4, 4, 3), recon_schem=3, lchardecomp=True, gamma=p.gamma,
                    mach=p.mach)
    dxi, jac = metrics(dims)
    dt = 0.5 / dims[0] / 2.0
    o = box.rk3_steps(st, dxi, jac, dt, nsteps=6, bctype=(0, 0, 0, 0, 1, 1), udf=1)
    inner = (slice(HM, -HM), slice(HM, -HM), HM)
    assert np.all(np.isfinite(o["q"][inner]))
    assert o["rho"][inner].min() > 0.1 and o["prs"][inner].min() > 0.02
    assert not np.array_equal(o["q"][inner], st["q"][inner])
