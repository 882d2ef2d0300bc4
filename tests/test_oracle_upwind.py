"""Pins of the checker's explicit upwind path (oracle/upwind.c).

The reference routines (src/flux.F90, src/riemann.F90, src/solver.F90
convrsduwd/chardecomp, src/commcal.F90 ducrossensor) sit in modules that
need MPI (flux.F90 uses the parallel module), which the image lacks, so
they cannot be compiled here; no reference test holds their outputs.
Each function is pinned by what the mathematics fixes instead, chosen so
that a dropped term, a wrong sign or a shifted stencil fails:
  * reconstructions: exactness on polynomials up to the design degree
    (and failure one degree above), MP = linear scheme on smooth data,
    no new extrema at a step, WENO weights -> linear weights on smooth data;
  * Steger-Warming: F+ + F- = F, pure upwinding for |M| >= 1, the sign of
    the split eigenvalues, and F+ and F- each against the textbook split;
  * ROUND: its normalised-variable formula at many points, the NVD corners,
    the linear-scheme limit and affine invariance;
  * chardecomp: LEV REV = I and REV diagonalises the Roe-averaged flux
    Jacobian along the normal;
  * convrsduwd: free-stream preservation, telescoping conservation on a
    periodic box, design-order convergence of the flux difference;
  * ducrossensor: closed forms (pure rotation -> 0, pure dilatation ->
    the pressure term) and the asymmetric -4..+5 window.
"""
import math

import numpy as np
import pytest

import pyoracle as po

HM = 5


def cell_avgs(coef, centers):
    """Unit-cell averages of the polynomial sum c_p x^p at the centers."""
    P = np.polynomial.Polynomial(coef).integ()
    return np.array([P(c + 0.5) - P(c - 0.5) for c in centers])


@pytest.mark.parametrize("name,npts,deg", [("suw3", 3, 2), ("suw5", 5, 4), ("suw7", 7, 6)])
def test_linear_upwind_exact_to_design_degree(name, npts, deg):
    # stencil centers i-(npts//2) .. i+npts//2 around i = 0; the value at the
    # interface x = +1/2 of the function whose cell averages are given
    centers = np.arange(npts) - npts // 2
    rng = np.random.default_rng(3)
    for d in range(deg + 1):
        c = np.zeros(d + 1)
        c[d] = 1.0
        c[:d] = rng.standard_normal(d)
        u = cell_avgs(c, centers)
        assert abs(po.recon(name, u) - np.polynomial.Polynomial(c)(0.5)) < 1e-11, d
    c = np.zeros(deg + 2)
    c[-1] = 1.0
    u = cell_avgs(c, centers)
    assert abs(po.recon(name, u) - np.polynomial.Polynomial(c)(0.5)) > 1e-6


def test_mp_equals_linear_scheme_on_smooth_monotone_data():
    x = np.arange(7) * 0.05
    u = np.exp(x)  # smooth, monotone, convex: the MP bounds are not active
    assert po.recon("mp7", u) == po.recon("suw7", u)
    assert po.recon("mp5", u[1:6]) == po.recon("suw5", u[1:6])


@pytest.mark.parametrize("name,n", [("mp7", 7), ("mp5", 5)])
def test_mp_adds_no_new_extremum_at_a_step(name, n):
    for k in range(1, n):
        u = np.where(np.arange(n) < k, 1.0, 0.0)
        v = po.recon(name, u)
        mid = n // 2
        lo, hi = min(u[mid], u[mid + 1]), max(u[mid], u[mid + 1])
        assert lo - 1e-14 <= v <= hi + 1e-14, (k, v)
    # the linear scheme does overshoot there, so the limiter is what acts
    u = np.where(np.arange(n) < n // 2, 1.0, 0.0)
    lin = po.recon("suw7" if n == 7 else "suw5", u)
    assert lin > 1.0 + 1e-3 or lin < -1e-3


def test_mp7ld_linear_weights_and_switch():
    # with the shock flag off MP7LD is its linear 8-point scheme; weight 0
    # reduces it to the 7th-order upwind one, weight 1 to the 8th-order
    # central one (coefficients of src/flux.F90:652-659)
    u = np.sin(np.arange(8) * 0.3)
    lin0 = po.recon("mp7ld", u, 0.0, 0)
    assert abs(lin0 - po.recon("suw7", u[:7])) < 1e-15
    c = np.polynomial.Polynomial(np.random.default_rng(5).standard_normal(8))
    u8 = cell_avgs(c.coef, np.arange(8) - 3)
    assert abs(po.recon("mp7ld", u8, 1.0, 0) - c(0.5)) < 1e-11
    step = np.where(np.arange(8) < 4, 1.0, 0.0)
    v = po.recon("mp7ld", step, 0.7, 1)
    assert -1e-14 <= v <= 1.0 + 1e-14


@pytest.mark.parametrize("name,n,lin", [("weno5", 5, "suw5"), ("weno5z", 5, "suw5"),
                                        ("weno7", 7, "suw7"), ("weno7z", 7, "suw7")])
def test_weno_tends_to_linear_scheme_on_smooth_data(name, n, lin):
    errs = []
    for h in (0.04, 0.02):
        u = np.sin(1.0 + h * (np.arange(n) - n // 2))
        errs.append(abs(po.recon(name, u) - po.recon(lin, u)))
    assert errs[1] < errs[0] and errs[1] < 1e-7
    assert abs(po.recon(name, np.full(n, 2.5)) - 2.5) < 1e-14


def round_nvd(z):
    """ROUND's face value in the normalised-variable diagram (Deng, "A
    unified framework for non-linear reconstruction schemes in a compact
    stencil", J. Comput. Phys. 2023), as a function of the normalised
    upwind value z = (u_i - u_{i-1}) / (u_{i+1} - u_{i-1}):
      g(z) = [p1 (1 - w1) + p2 w1] (1 - w2) + p3 w2,
    p1 the third-order linear scheme 5z/6 + 1/3 with the two
    boundedness corrections max(0, 1100 (z-0.05)^3 (0.47-z)^3) and
    max(0, 18000 (z-0.55)^3 (0.97-z)^5), p2 = 3z/2 near z = 0, p3 = (z+1)/2
    near z = 1, weights w1 = (1 + 12 z^2)^-4, w2 = (1 + 5 (z-1)^2)^-8; the
    face value is u_{i-1} + g(z) (u_{i+1} - u_{i-1})."""
    bump = (max(0.0, 1100.0 * (z - 0.05) ** 3 * (0.47 - z) ** 3)
            + max(0.0, 18000.0 * (z - 0.55) ** 3 * (0.97 - z) ** 5))
    p1 = 5.0 / 6.0 * z + 1.0 / 3.0 + bump
    w1 = (1.0 + 12.0 * z * z) ** -4
    w2 = (1.0 + 5.0 * (z - 1.0) ** 2) ** -8
    return (p1 * (1.0 - w1) + 1.5 * z * w1) * (1.0 - w2) + (0.5 * z + 0.5) * w2


def test_round_preserves_constants_and_linear_data():
    # on a linear profile (z = 1/2) ROUND is the third-order linear scheme:
    # -u0/6 + 5 u1/6 + u2/3
    for u in ([1.0, 2.0, 3.0], [-4.0, -1.0, 2.0], [7.5, 7.25, 7.0]):
        u = np.array(u)
        lin = -u[0] / 6 + 5 * u[1] / 6 + u[2] / 3
        assert abs(po.recon("round", u) - lin) <= 1e-14 * max(1.0, abs(lin)), u
    assert po.recon("round", np.full(3, 2.5)) == 2.5
    v = po.recon("round", np.array([0.0, 0.0, 1.0]))
    assert 0.0 <= v <= 1e-6


def test_round_against_its_normalised_variable_formula():
    """At 2000 random triples (z from -3 to 4 and across [0, 1] densely):
    the face value is u_{i-1} + g(z) (u_{i+1} - u_{i-1}) with g the NVD
    formula; g passes through the NVD corners (0, 0) and (1, 1) and the
    linear-scheme point (1/2, 3/4); g(z) >= z on [0, 1] (never more
    diffusive than first-order upwind there); and the face value moves with
    any affine change u -> a u + b of the data (a > 0)."""
    rng = np.random.default_rng(5)
    worst = 0.0
    for k in range(2000):
        z = rng.uniform(-3.0, 4.0) if k % 2 else rng.uniform(0.0, 1.0)
        u0 = rng.standard_normal()
        du = rng.choice([-1.0, 1.0]) * (0.1 + rng.random())
        u = np.array([u0, u0 + z * du, u0 + du])
        got = po.recon("round", u)
        zz = (u[1] - u[0] + 1e-16) / (u[2] - u[0] + 1e-16)
        want = round_nvd(zz) * (u[2] - u[0]) + u[0]
        worst = max(worst, abs(got - want) / (abs(du) + abs(u0)))
        a, b = 0.5 + 3.0 * rng.random(), rng.standard_normal()
        ga = po.recon("round", a * u + b)
        assert abs(ga - (a * got + b)) <= 1e-12 * (abs(a * got) + abs(b) + a * abs(du))
    assert worst <= 1e-14, worst
    assert round_nvd(1.0) == 1.0 and abs(round_nvd(0.0)) < 1e-6
    assert abs(round_nvd(0.5) - 0.75) < 1e-15
    zs = np.linspace(0.0, 1.0, 2001)
    assert all(round_nvd(z) >= z - 1e-15 for z in zs)
    # slope of the linear scheme at z = 1/2 (third order there)
    h = 1e-4
    assert abs((round_nvd(0.5 + h) - round_nvd(0.5 - h)) / (2 * h) - 5.0 / 6.0) < 5e-3


def test_recons_exp_boundary_degradation():
    f = np.arange(8.0) ** 2
    assert po.recons_exp(f, 0, 40, 1, 3) == 0.5 * (f[3] + f[4])
    assert po.recons_exp(f, 39, 40, 2, 3) == 0.5 * (f[3] + f[4])
    assert po.recons_exp(f, 1, 40, 1, 3) == po.recon("suw3", f[2:5])
    assert po.recons_exp(f, 2, 40, 1, 0) == po.recon("suw5", f[1:6])
    assert po.recons_exp(f, 10, 40, 1, 0) == po.recon("suw7", f[:7])
    assert po.recons_exp(f, 10, 40, 3, -1) == f[3]
    # a periodic line is never degraded
    assert po.recons_exp(f, 0, 40, 3, 0) == po.recon("suw7", f[:7])


@pytest.mark.parametrize("seed", range(6))
def test_steger_warming_split_sums_to_the_flux(seed):
    rng = np.random.default_rng(seed)
    gamma, mach = 1.4, 1.0
    rho = 0.5 + rng.random()
    T = 0.5 + rng.random()
    p = rho * T / (gamma * mach ** 2)
    c = math.sqrt(T) / mach
    dxi = rng.standard_normal(3)
    nrm = np.linalg.norm(dxi)
    for mn in (-2.0, -0.5, 0.0, 0.3, 0.999, 1.5):
        n = dxi / nrm
        vel = mn * c * n + 0.3 * np.cross(n, rng.standard_normal(3))
        E = p / (gamma - 1) + 0.5 * rho * vel @ vel
        q = np.array([rho, *(rho * vel), E])
        J = 0.3 + rng.random()
        fp, fm = po.sw_split(rho, vel, p, T, q, dxi, J, gamma, mach)
        U = dxi @ vel
        F = J * np.array([rho * U, *(rho * vel * U + dxi * p), (E + p) * U])
        assert np.allclose(fp + fm, F, rtol=1e-13, atol=1e-13 * np.abs(F).max()), mn
        if mn >= 1.0:
            assert np.all(fm == 0.0)
            assert np.allclose(fp, F, rtol=1e-14, atol=1e-14 * np.abs(F).max())
        if mn <= -1.0:
            assert np.all(fp == 0.0)
        # split mass fluxes carry the sign of their eigenvalues
        assert fp[0] >= -1e-15 and fm[0] <= 1e-15


def textbook_sw(rho, vel, p, c, dxi, J, gamma, eps=0.04):
    """The Steger-Warming split of the flux along grad(xi) (Steger & Warming,
    J. Comput. Phys. 40 (1981); Toro, Riemann Solvers, sec. 8.4), written
    in contravariant form: eigenvalues U, U, U, U + c|grad xi|, U -
    c|grad xi| with U = grad(xi).u, each split as (l +- sqrt(l^2 + eps^2))/2
    (the smoothed split with eps = 0.04 that src/riemann.F90:23-161 uses),
    unit normal n = grad(xi)/|grad xi|, un = n.u, a the speed of sound:
      F+- = J rho/(2 gamma) [ al,
                              al u + a n (l3 - l1),
                              al |u|^2/2 + a un (l3 - l1) + a^2 (l1 + l3)/(gamma - 1) ]
    with al = 2 (gamma - 1) l2 + l1 + l3, l1 = U - c|grad xi|, l2 = U,
    l3 = U + c|grad xi| (their split parts)."""
    g = np.linalg.norm(dxi)
    n = dxi / g
    U = dxi @ vel
    un = n @ vel
    lam = np.array([U - c * g, U, U + c * g])
    out = []
    for sgn in (1.0, -1.0):
        ls = 0.5 * (lam + sgn * np.sqrt(lam * lam + eps * eps))
        l1, l2, l3 = ls
        al = 2.0 * (gamma - 1.0) * l2 + l1 + l3
        f = np.empty(5)
        f[0] = al
        f[1:4] = al * vel + c * n * (l3 - l1)
        f[4] = al * 0.5 * (vel @ vel) + c * un * (l3 - l1) + c * c * (l1 + l3) / (gamma - 1.0)
        out.append(J * rho / (2.0 * gamma) * f)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_steger_warming_halves_match_the_textbook_split(seed):
    """F+ and F- each (not only their sum) against the textbook split of the
    flux along grad(xi) with the eps-smoothed eigenvalues, at subsonic normal
    Mach numbers of either sign; above |M| = 1 the whole flux goes upwind."""
    rng = np.random.default_rng(100 + seed)
    gamma, mach = 1.4, 1.0
    rho = 0.5 + rng.random()
    T = 0.5 + rng.random()
    p = rho * T / (gamma * mach ** 2)
    c = math.sqrt(T) / mach
    dxi = rng.standard_normal(3) * (0.5 + 2.0 * rng.random())
    nrm = np.linalg.norm(dxi)
    n = dxi / nrm
    for mn in (-0.95, -0.5, -0.01, 0.0, 0.02, 0.4, 0.9):
        vel = mn * c * n + 0.4 * np.cross(n, rng.standard_normal(3))
        E = p / (gamma - 1) + 0.5 * rho * vel @ vel
        q = np.array([rho, *(rho * vel), E])
        J = 0.3 + rng.random()
        fp, fm = po.sw_split(rho, vel, p, T, q, dxi, J, gamma, mach)
        tp, tm = textbook_sw(rho, vel, p, c, dxi, J, gamma)
        scale = max(np.abs(tp).max(), np.abs(tm).max())
        assert np.allclose(fp, tp, rtol=0, atol=1e-13 * scale), (mn, fp, tp)
        assert np.allclose(fm, tm, rtol=0, atol=1e-13 * scale), (mn, fm, tm)
        # a plausible slip (the acoustic eigenvalues swapped, or the
        # smoothing dropped) is far outside that bound
        wrong, _ = textbook_sw(rho, vel, p, c, dxi, J, gamma, eps=0.0)
        assert np.abs(wrong - tp).max() > 1e-6 * scale


def roe_jacobian(R_state, n, gamma):
    """Flux Jacobian along unit normal n at a state (rho, u, H)."""
    u, H = R_state
    un = u @ n
    k = 0.5 * u @ u
    g1 = gamma - 1.0
    A = np.zeros((5, 5))
    A[0, 1:4] = n
    for a in range(3):
        A[1 + a, 0] = g1 * k * n[a] - u[a] * un
        A[1 + a, 1:4] = u[a] * n - g1 * u * n[a]
        A[1 + a, 1 + a] += un
        A[1 + a, 4] = g1 * n[a]
    A[4, 0] = (g1 * k - H) * un
    A[4, 1:4] = H * n - g1 * u * un
    A[4, 4] = gamma * un
    return A


@pytest.mark.parametrize("axis", [0, 1, 2, None])
def test_chardecomp_inverse_pair_and_diagonalises_roe_jacobian(axis):
    rng = np.random.default_rng(11 + (axis or 0))
    gamma = 1.4
    states = []
    for _ in range(2):
        rho = 0.5 + rng.random()
        vel = rng.standard_normal(3)
        p = 0.5 + rng.random()
        E = p / (gamma - 1) + 0.5 * rho * vel @ vel
        ddi = np.zeros(3)
        if axis is None:
            ddi = rng.standard_normal(3)
        else:
            ddi[axis] = 1.0 + rng.random()
        states.append({"tmp": 1.0, "rho": rho, "prs": p, "E": E, "vel": vel, "ddi": ddi})
    R, L = po.chardecomp(states[0], states[1], gamma)
    assert np.allclose(L @ R, np.eye(5), atol=1e-12)
    # Roe average
    sl, sr = math.sqrt(states[0]["rho"]), math.sqrt(states[1]["rho"])
    w = sl / (sl + sr)
    u = w * states[0]["vel"] + (1 - w) * states[1]["vel"]
    H = (w * (states[0]["E"] + states[0]["prs"]) / states[0]["rho"] +
         (1 - w) * (states[1]["E"] + states[1]["prs"]) / states[1]["rho"])
    nrm = 0.5 * (states[0]["ddi"] + states[1]["ddi"])
    n = nrm / np.linalg.norm(nrm)
    A = roe_jacobian((u, H), n, gamma)
    c = math.sqrt((gamma - 1) * (H - 0.5 * u @ u))
    lam = np.array([u @ n - c, u @ n, u @ n, u @ n, u @ n + c])
    assert np.allclose(L @ A @ R, np.diag(lam), atol=1e-11 * max(1.0, np.abs(lam).max()))


def test_iwind8_stencils_and_clamp():
    assert [po.uwd_lib().orc_iwind8(10, n, -5, 45, 1) for n in range(1, 9)] == list(range(7, 15))
    assert [po.uwd_lib().orc_iwind8(10, n, -5, 45, 0) for n in range(1, 9)] == list(
        range(14, 6, -1))
    assert po.uwd_lib().orc_iwind8(1, 1, 0, 40, 1) == 0
    assert po.uwd_lib().orc_iwind8(39, 1, 0, 40, 0) == 40


# ------------------------------------------------------------- box level

def box_metrics(dims, lengths):
    im, jm, km = dims
    hs = (im + 1 + 2 * HM, jm + 1 + 2 * HM, km + 1 + 2 * HM)
    dxi = np.zeros(hs + (9,), order="F")
    h = []
    for d, n in enumerate(dims):
        if n == 0:
            h.append(1.0)
            dxi[..., d + 3 * d] = 1.0
            continue
        h.append(lengths[d] / n)
        dxi[..., d + 3 * d] = 1.0 / h[-1]
    jac = np.full(hs, h[0] * h[1] * h[2], order="F")
    return dxi, jac, h


def coords(dims, lengths):
    """Node coordinates on -hm..n+hm, taken periodically (node t and node
    t mod n get bit-identical coordinates, so periodic fields are exactly
    periodic, halos included)."""
    out = []
    for d, n in enumerate(dims):
        h = lengths[d] / n if n else 0.0
        t = np.arange(-HM, n + HM + 1)
        out.append((np.mod(t, n) if n else 0 * t) * h)
    return np.meshgrid(*out, indexing="ij")


@pytest.mark.parametrize("recon,lchar", [(3, True), (0, False), (1, True), (6, False)])
def test_free_stream_is_preserved(recon, lchar):
    dims = (24, 20, 16)
    box = po.UwdBox(dims, recon_schem=recon, lchardecomp=lchar)
    X = coords(dims, (1, 1, 1))
    one = np.ones_like(X[0])
    st = po.prim_to_state(1.3 * one, 0.4 * one, -0.2 * one, 0.1 * one, 0.9 * one)
    dxi, jac, _ = box_metrics(dims, (1, 1, 1))
    r = box.rhscal(st, dxi, jac)
    assert np.all(r["qrhs"] == 0.0)


def periodic_wave(dims, lengths, mach=1.0):
    X = coords(dims, lengths)
    tw = [2 * np.pi / L for L in lengths]
    rho = 1.0 + 0.2 * np.sin(tw[0] * X[0]) * np.cos(tw[1] * X[1]) + 0.1 * np.sin(tw[2] * X[2])
    u = 0.3 + 0.1 * np.cos(tw[1] * X[1])
    v = -0.2 + 0.05 * np.sin(tw[0] * X[0])
    w = 0.1 + 0.05 * np.cos(tw[2] * X[2])
    p = 1.0 / 1.4 + 0.05 * np.cos(tw[0] * X[0] + tw[2] * X[2])
    return po.prim_to_state(rho, u, v, w, p, mach=mach)


@pytest.mark.parametrize("recon,lchar", [(3, True), (0, False), (2, True)])
def test_periodic_box_conserves(recon, lchar):
    dims = (24, 20, 16)
    lengths = (1.0, 1.0, 1.0)
    box = po.UwdBox(dims, recon_schem=recon, lchardecomp=lchar, shkcrt=-1.0)
    st = periodic_wave(dims, lengths)
    dxi, jac, _ = box_metrics(dims, lengths)
    r = box.rhscal(st, dxi, jac)["qrhs"]
    # unique points 0..n-1 on each periodic axis: the flux differences
    # telescope to zero
    tot = r[:-1, :-1, :-1].sum(axis=(0, 1, 2))
    scale = np.abs(r).max()
    assert np.all(np.abs(tot) < 1e-11 * scale * r[:-1, :-1, :-1].shape[0] ** 3), tot


def flux_x(st, dxi_x, jac):
    rho, prs = st["rho"], st["prs"]
    q = st["q"]
    u = st["vel"][..., 0]
    U = dxi_x * u
    return np.stack([jac * q[..., 0] * U, jac * (q[..., 1] * U + dxi_x * prs),
                     jac * q[..., 2] * U, jac * q[..., 3] * U,
                     jac * (q[..., 4] + prs) * U], axis=-1)


def test_x_flux_difference_converges_at_design_order():
    """A 1-D wave in x (periodic box): -qrhs from convrsduwd with SUW7 and no
    characteristic projection approximates dF/dxi with O(h^7) error."""
    errs = []
    for n in (24, 48):
        dims = (n, 16, 16)
        lengths = (1.0, 1.0, 1.0)
        X = coords(dims, lengths)
        kx = 2 * np.pi
        rho = 1.0 + 0.2 * np.sin(kx * X[0])
        u = 0.5 + 0.1 * np.cos(kx * X[0])
        p = 1.0 / 1.4 + 0.05 * np.sin(kx * X[0])
        st = po.prim_to_state(rho, u, 0 * u, 0 * u, p)
        dxi, jac, h = box_metrics(dims, lengths)
        box = po.UwdBox(dims, recon_schem=0, lchardecomp=False)
        q = box.convrsduwd(st, dxi, jac)
        # the mass flux F = J rho u / h: its index-space derivative is
        # J d(rho u)/dx, in closed form
        xs = np.arange(0, n + 1) * h[0]
        J = h[0] * h[1] * h[2]
        exact = J / h[0] * h[0] * (0.2 * kx * np.cos(kx * xs) * (0.5 + 0.1 * np.cos(kx * xs)) +
                                   (1.0 + 0.2 * np.sin(kx * xs)) * (-0.1 * kx * np.sin(kx * xs)))
        err = np.abs(q[:, 5, 5, 0] - exact).max() / np.abs(exact).max()
        errs.append(err)
    rate = math.log2(errs[0] / errs[1])
    assert rate > 6.0, (errs, rate)


def test_ducros_closed_forms_and_window():
    dims = (20, 18, 16)
    box = po.UwdBox(dims, npdc=(3, 3, 3), shkcrt=0.05)
    hs = box.hshape
    ish = box.ishape
    prs = np.ones(hs, order="F")
    # pure rotation: no dilatation -> ssf = 0
    dvel = np.zeros(ish + (9,), order="F")
    dvel[..., 1 + 3 * 0] = 1.0   # dv/dx
    dvel[..., 0 + 3 * 1] = -1.0  # du/dy
    ssf, lsh = box.ducros(dvel, prs)
    assert np.all(ssf == 0.0) and not lsh.any()
    # pure dilatation with one pressure kink at (7, 9, 8): ssf there is the
    # pressure ratio; the flag covers i-5..i+4 along each axis only
    dvel[:] = 0.0
    dvel[..., 0] = 1.0  # du/dx
    prs[7 + HM, 9 + HM, 8 + HM] = 2.0
    ssf, lsh = box.ducros(dvel, prs)
    centre = ssf[7 + HM, 9 + HM, 8 + HM]
    assert abs(centre - 1.0 / (1.0 + 1e-30) * (2.0 / 6.0)) < 1e-15
    assert abs(ssf[8 + HM, 9 + HM, 8 + HM] - 1.0 / 5.0) < 1e-15
    flagged = set(zip(*np.nonzero(lsh)))
    expect = set()
    for c in [(7, 9, 8), (6, 9, 8), (8, 9, 8), (7, 8, 8), (7, 10, 8), (7, 9, 7), (7, 9, 9)]:
        for d in range(3):
            for o in range(-5, 5):
                p = list(c)
                p[d] += o
                if 0 <= p[d] <= dims[d]:
                    expect.add(tuple(p))
    assert flagged == expect


def test_undefined_closures_are_rejected():
    dims = (16, 16, 16)
    box = po.UwdBox(dims, npdc=(3, 5, 3), recon_schem=0, lchardecomp=False)
    st = periodic_wave(dims, (1, 1, 1))
    dxi, jac, _ = box_metrics(dims, (1, 1, 1))
    with pytest.raises(ValueError):
        box.convrsduwd(st, dxi, jac)


@pytest.mark.parametrize("npdc", [(1, 2, 3), (2, 1, 3), (4, 3, 3), (3, 3, 1)])
def test_bounded_boxes_run_and_keep_uniform_flow_nearly_steady(npdc):
    dims = (20, 18, 16)
    box = po.UwdBox(dims, npdc=npdc, recon_schem=3, lchardecomp=True)
    X = coords(dims, (1, 1, 1))
    one = np.ones_like(X[0])
    st = po.prim_to_state(1.1 * one, 0.3 * one, 0.2 * one, -0.1 * one, 0.8 * one)
    dxi, jac, _ = box_metrics(dims, (1, 1, 1))
    r = box.rhscal(st, dxi, jac)["qrhs"]
    assert np.abs(r).max() < 1e-13 * np.abs(flux_x(st, dxi[..., 0], jac)).max()


def test_two_dimensional_box():
    dims = (24, 20, 0)
    box = po.UwdBox(dims, npdc=(3, 3, 3), recon_schem=3, lchardecomp=True)
    X = coords(dims, (1, 1, 1))
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X[0]) * np.cos(2 * np.pi * X[1])
    one = np.ones_like(rho)
    st = po.prim_to_state(rho, 0.3 * one, -0.2 * one, 0 * one, 0.7 * one)
    dxi, jac, _ = box_metrics(dims, (1, 1, 1))
    r = box.rhscal(st, dxi, jac)
    q = r["qrhs"]
    assert q.shape == (25, 21, 1, 5)
    assert np.all(q[..., 3] == 0.0)  # no z momentum flux in 2-D
    tot = q[:-1, :-1, 0].sum(axis=(0, 1))
    assert np.all(np.abs(tot) < 1e-12 * np.abs(q).max() * 500)


# ------------------------------------------------ compact upwind ('543c')

def line_of_cell_avgs(H, dim, lo=-HM):
    """f(-hm:dim+hm) = unit-cell averages of H' (H an antiderivative)."""
    t = np.arange(lo, dim + HM + 1)
    return H(t + 0.5) - H(t - 0.5)


@pytest.mark.parametrize("ntype", [1, 2, 3, 4])
@pytest.mark.parametrize("plus", [True, False])
def test_flux_compact_preserves_constants(ntype, plus):
    dim = 40
    fh = po.flux_compact(ntype, np.full(dim + 1 + 2 * HM, 1.7), plus)
    assert np.allclose(fh, 1.7, rtol=0, atol=1e-14)


@pytest.mark.parametrize("plus", [True, False])
def test_flux_compact_exact_on_polynomials(plus):
    """Every row of the inner (ntype 3) system is exact for polynomials of
    degree <= 4 (5th-order compact upwind interior, 6th-order explicit
    interface rows), so the solve returns the interface values h(j+1/2)
    exactly; the 3rd-order boundary rows (ntype 1, 4) keep degree <= 2.
    One degree higher is not reproduced."""
    dim = 30
    rng = np.random.default_rng(9)
    for ntype, deg in ((3, 4), (1, 2), (4, 2), (2, 2)):
        for d in range(deg + 2):
            c = np.zeros(d + 1)
            c[d] = 1.0
            c[:d] = rng.standard_normal(d)
            h = np.polynomial.Polynomial(c) * (1.0 / dim ** d)  # keep values O(1)
            H = h.integ()
            fh = po.flux_compact(ntype, line_of_cell_avgs(H, dim), plus, bfac=0.7)
            exact = h(np.arange(-1, dim + 1) + 0.5)
            err = np.abs(fh - exact).max()
            if d <= deg:
                assert err < 1e-11, (ntype, d, err)
            elif ntype == 3:
                assert err > 1e-9, (ntype, d, err)


def test_flux_compact_converges_at_fifth_order():
    errs = []
    for dim in (32, 64):
        h = 1.0 / dim
        t = np.arange(-HM, dim + HM + 1)
        k = 2 * np.pi
        f = (np.cos(k * (t - 0.5) * h) - np.cos(k * (t + 0.5) * h)) / (k * h)
        fh = po.flux_compact(3, f, True, bfac=0.7)
        exact = np.sin(k * (np.arange(-1, dim + 1) + 0.5) * h)
        errs.append(np.abs(fh - exact)[5:-5].max())
    assert math.log2(errs[0] / errs[1]) > 4.5, errs


def test_mplimiter_boundary_and_interior():
    f = np.array([1.0, 1.0, 1.0, 0.0, 0.0])
    fl = 1.3  # an overshooting linear value
    assert po.mplimiter(f, fl, True, 0, 40, 1) == fl
    assert po.mplimiter(f, fl, True, 38, 40, 2) == fl
    v = po.mplimiter(f, fl, True, 10, 40, 3)
    assert 0.0 <= v <= 1.0
    assert po.mplimiter(f, fl, False, 10, 40, 3) == fl  # shock flag off


@pytest.mark.parametrize("lchar", [True, False])
def test_compact_upwind_free_stream(lchar):
    # (no telescoping identity here: the non-cyclic ntype-3 compact systems
    # of periodic lines make Fh(-1) and Fh(n-1) differ at truncation level)
    dims = (24, 20, 16)
    box = po.UwdBox(dims, recon_schem=3, lchardecomp=lchar, shkcrt=-1.0)
    X = coords(dims, (1, 1, 1))
    one = np.ones_like(X[0])
    st = po.prim_to_state(1.3 * one, 0.4 * one, -0.2 * one, 0.1 * one, 0.9 * one)
    dxi, jac, _ = box_metrics(dims, (1, 1, 1))
    r = box.rhscal(st, dxi, jac, compact=True)
    assert np.abs(r["qrhs"]).max() < 1e-13


def test_compact_upwind_converges_at_design_order():
    errs = []
    for n in (24, 48):
        dims = (n, 16, 16)
        X = coords(dims, (1, 1, 1))
        kx = 2 * np.pi
        rho = 1.0 + 0.2 * np.sin(kx * X[0])
        u = 0.5 + 0.1 * np.cos(kx * X[0])
        p = 1.0 / 1.4 + 0.05 * np.sin(kx * X[0])
        st = po.prim_to_state(rho, u, 0 * u, 0 * u, p)
        dxi, jac, h = box_metrics(dims, (1, 1, 1))
        box = po.UwdBox(dims, recon_schem=3, lchardecomp=False)
        q = box.rhscal(st, dxi, jac, compact=True)["qrhs"]
        xs = np.arange(0, n + 1) * h[0]
        J = h[0] * h[1] * h[2]
        exact = -J * (0.2 * kx * np.cos(kx * xs) * (0.5 + 0.1 * np.cos(kx * xs)) +
                      (1.0 + 0.2 * np.sin(kx * xs)) * (-0.1 * kx * np.sin(kx * xs)))
        errs.append(np.abs(q[:, 5, 5, 0] - exact).max() / np.abs(exact).max())
    assert math.log2(errs[0] / errs[1]) > 4.5, errs
