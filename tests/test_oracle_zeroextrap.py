"""The checker's zeroextrap (bctype 50, src/bc.F90:2331-2627) and symmetry
(bctype 60, src/bc.F90:2639-2971).  zeroextrap: the boundary
node of a face from its two inner neighbours by extrapolate_2o with dv = 0
(src/commfunc.F90:277-285), v0 = (4 v1 - v2)/3.  That is the value at the
face of the parabola with zero slope there, so a field c + d r^2 (r the
node distance from the face) comes back as c on the face, which a linear
extrapolation (c + 2d/3 for it), a wrong weight or the wrong neighbour
fails.  Faces 1-2 extrapolate the temperature and take the density from
the equation of state, faces 3-6 the other way round.  symmetry: the same
extrapolation with the velocity normal to the face zero and the others
kept, on a rectilinear grid."""
import numpy as np
import pytest

import pyoracle as po

HM = 5
GAMMA, MACH = 1.4, 0.4


def parabolic_state(dims, face, c, d):
    """Primitives c[k] + d[k] r^2 in the distance r (in nodes) from the given
    face, constant along the face; every other node keeps that law too."""
    shape = tuple(n + 1 + 2 * HM for n in dims)
    ax, hi = (face - 1) // 2, (face - 1) % 2
    idx = np.arange(shape[ax]) - HM
    r = (dims[ax] - idx) if hi else idx
    sh = [1, 1, 1]
    sh[ax] = shape[ax]
    r = r.reshape(sh).astype(np.float64)
    prim = [np.broadcast_to(c[k] + d[k] * r * r, shape).copy() for k in range(5)]
    return po.prim_to_state(*prim, gamma=GAMMA, mach=MACH), r


@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_zeroextrap_returns_the_zero_slope_parabola(face):
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, face, c, d)
    o = box.zeroextrap(st, face)
    ax, hi = (face - 1) // 2, (face - 1) % 2
    sl = [slice(HM, HM + n + 1) for n in dims]
    sl[ax] = HM + (dims[ax] if hi else 0)
    sl = tuple(sl)
    const2 = GAMMA * MACH * MACH
    u = [o["vel"][sl + (m,)] for m in range(3)]
    p, rho, t = o["prs"][sl], o["rho"][sl], o["tmp"][sl]
    for m in range(3):
        assert np.max(np.abs(u[m] - c[1 + m])) <= 1e-14, (face, m)
    assert np.max(np.abs(p - c[4])) <= 1e-14 * c[4]
    if face <= 2:
        # the temperature extrapolated (its own law is not a parabola in r,
        # so only the equation of state is checked), the density from it
        assert np.max(np.abs(rho - p / t * const2)) <= 1e-14
    else:
        assert np.max(np.abs(rho - c[0])) <= 1e-14
        assert np.max(np.abs(t - p / rho * const2)) <= 1e-14
    q = o["q"][sl]
    assert np.array_equal(q[..., 0], rho)
    for m in range(3):
        assert np.max(np.abs(q[..., 1 + m] - rho * u[m])) <= 1e-15
    e = p / (GAMMA - 1.0) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    assert np.max(np.abs(q[..., 4] - e)) <= 1e-13 * np.max(np.abs(e))
    # every other node untouched
    mask = np.ones(st["prs"].shape, dtype=bool)
    mask[sl] = False
    for nm in ("prs", "rho", "tmp"):
        assert np.array_equal(o[nm][mask], st[nm][mask]), nm


def test_zeroextrap_is_not_a_linear_extrapolation():
    """A linear law c + b r is extrapolated to c + 2b/3, not c."""
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    shape = tuple(n + 1 + 2 * HM for n in dims)
    r = (np.arange(shape[0]) - HM).reshape(-1, 1, 1).astype(np.float64) * np.ones(shape)
    st = po.prim_to_state(np.full(shape, 1.0), 0.2 + 0.03 * r, np.zeros(shape), np.zeros(shape),
                          np.full(shape, 1.0 / (GAMMA * MACH * MACH)), gamma=GAMMA, mach=MACH)
    o = box.zeroextrap(st, 1)
    u0 = o["vel"][HM, HM:HM + dims[1] + 1, HM, 0]
    assert np.max(np.abs(u0 - (0.2 + 0.02))) <= 1e-15


def test_central_box_accepts_zeroextrap():
    ob = po.OracleBox((16, 12, 10), variant=0, npdc=(4, 4, 3))
    ob.set_bc((50, 50, 50, 50, 0, 0))
    with pytest.raises(ValueError):
        ob.set_bc((51, 0, 0, 0, 0, 0))


@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_symmetry_zeroes_the_normal_velocity(face):
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, face, c, d)
    o = box.symmetry(st, face)
    ax, hi = (face - 1) // 2, (face - 1) % 2
    sl = [slice(HM, HM + n + 1) for n in dims]
    sl[ax] = HM + (dims[ax] if hi else 0)
    sl = tuple(sl)
    const2 = GAMMA * MACH * MACH
    u = [o["vel"][sl + (m,)] for m in range(3)]
    p, rho, t = o["prs"][sl], o["rho"][sl], o["tmp"][sl]
    for m in range(3):
        want = 0.0 if m == ax else c[1 + m]
        assert np.max(np.abs(u[m] - want)) <= 1e-14, (face, m)
    assert np.max(np.abs(p - c[4])) <= 1e-14 * c[4]
    assert np.max(np.abs(rho - p / t * const2)) <= 1e-14
    q = o["q"][sl]
    assert np.array_equal(q[..., 0], rho)
    assert np.all(q[..., 1 + ax] == 0.0)
    # both of fvar2q's energy forms give the same value to round-off
    e = p / (GAMMA - 1.0) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    assert np.max(np.abs(q[..., 4] - e)) <= 1e-13 * np.max(np.abs(e))


def test_central_box_accepts_symmetry():
    ob = po.OracleBox((16, 12, 10), variant=0, npdc=(4, 4, 3))
    ob.set_bc((60, 60, 50, 60, 0, 0))


@pytest.mark.parametrize("face,bct", [(1, 41), (2, 42), (3, 41), (4, 42)])
def test_walls_hold_the_fluid_at_rest(face, bct):
    """noslip (41) and noslip_adibatic (42): the wall node at rest, its
    pressure the zero-slope parabola's face value, its temperature the face's
    wall temperature (41) or extrapolated alike (42), the density from the
    equation of state and the energy p/(gamma-1)."""
    dims = (12, 10, 0)
    tw = (1.3, 1.1, 0.9, 1.2, 1.0, 1.0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH, twall=tw)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, face, c, d)
    o = box.wall(st, face, bct == 41)
    ax, hi = (face - 1) // 2, (face - 1) % 2
    sl = [slice(HM, HM + n + 1) for n in dims]
    sl[ax] = HM + (dims[ax] if hi else 0)
    sl = tuple(sl)
    const2 = GAMMA * MACH * MACH
    for m in range(3):
        assert np.all(o["vel"][sl + (m,)] == 0.0)
    p, rho, t = o["prs"][sl], o["rho"][sl], o["tmp"][sl]
    assert np.max(np.abs(p - c[4])) <= 1e-14 * c[4]
    if bct == 41:
        assert np.all(t == tw[face - 1])
    else:
        # the extrapolated temperature of the parabolic state's own T
        t1 = st["tmp"][tuple(HM + (dims[ax] - 1 if hi else 1) if k == ax else s_
                             for k, s_ in enumerate(sl))]
        t2 = st["tmp"][tuple(HM + (dims[ax] - 2 if hi else 2) if k == ax else s_
                             for k, s_ in enumerate(sl))]
        assert np.max(np.abs(t - (4.0 * t1 - t2) / 3.0)) <= 1e-14
    assert np.max(np.abs(rho - p / t * const2)) <= 1e-14
    q = o["q"][sl]
    assert np.array_equal(q[..., 0], rho)
    assert np.all(q[..., 1:4] == 0.0)
    assert np.max(np.abs(q[..., 4] - p / (GAMMA - 1.0))) <= 1e-14 * np.max(p)


def test_walls_only_on_faces_1_to_4():
    ob = po.OracleBox((16, 12, 10), variant=0, npdc=(4, 4, 4))
    ob.set_bc((41, 42, 41, 42, 0, 0))
    with pytest.raises(ValueError):
        ob.set_bc((0, 0, 0, 0, 41, 0))


def test_outflow_face2_copies_the_inner_node():
    """outflow (21) on face 2: the zero-order branch the reference always takes
    (velocity, pressure and temperature of node n-1, the density from the
    equation of state)."""
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, 2, c, d)
    o = box.outflow(st, 2, 0.01)
    b, i1 = HM + dims[0], HM + dims[0] - 1
    J = slice(HM, HM + dims[1] + 1)
    for nm in ("prs", "tmp"):
        assert np.array_equal(o[nm][b, J, HM], st[nm][i1, J, HM]), nm
    for m in range(3):
        assert np.array_equal(o["vel"][b, J, HM, m], st["vel"][i1, J, HM, m])
    const2 = GAMMA * MACH * MACH
    assert np.max(np.abs(o["rho"][b, J, HM] - o["prs"][b, J, HM] / o["tmp"][b, J, HM] * const2)) \
        <= 1e-14


def _face4(dims, v0):
    """A state whose face-4 node (j = jm) holds the velocity v0 normal to the
    face while the inner nodes hold the zero-slope parabola."""
    c = (1.1, 0.3, 0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, 4, c, d)
    b = HM + dims[1]
    st["vel"][:, b, :, 1] = v0
    return st, c, b


def test_outflow_face4_supersonic_takes_the_extrapolation():
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    st, c, b = _face4(dims, 3.0)  # sqrt(T)/Ma is about 2.4 here
    o = box.outflow(st, 4, 0.01)
    I = slice(HM, HM + dims[0] + 1)
    assert np.max(np.abs(o["prs"][I, b, HM] - c[4])) <= 1e-14 * c[4]
    assert np.max(np.abs(o["rho"][I, b, HM] - c[0])) <= 1e-14
    assert np.max(np.abs(o["vel"][I, b, HM, 1] - c[2])) <= 1e-14
    const2 = GAMMA * MACH * MACH
    assert np.max(np.abs(o["tmp"][I, b, HM] - o["prs"][I, b, HM] / o["rho"][I, b, HM] * const2)) \
        <= 1e-14


def test_outflow_face4_subsonic_relaxes_towards_pinf():
    """Below the sound speed the new pressure keeps the outgoing characteristic
    p - rho c v of the node (dt = 0: p_new - rho c ve = p - rho c v, with the
    node's own rho, c = sqrt(T)/Ma and v), and relaxes towards pinf as dt
    grows (dt -> infinity: p_new -> pinf)."""
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    st, c, b = _face4(dims, 0.1)
    I = slice(HM, HM + dims[0] + 1)
    p0, r0, t0 = st["prs"][I, b, HM], st["rho"][I, b, HM], st["tmp"][I, b, HM]
    css = np.sqrt(t0) / MACH
    o = box.outflow(st, 4, 0.0)
    ve = o["vel"][I, b, HM, 1]
    assert np.max(np.abs((o["prs"][I, b, HM] - r0 * css * ve) - (p0 - r0 * css * 0.1))) \
        <= 1e-13 * np.max(p0)
    o = box.outflow(st, 4, 1e12)
    pinf = 1.0 / (GAMMA * MACH * MACH)
    assert np.max(np.abs(o["prs"][I, b, HM] - pinf)) <= 1e-9 * pinf


def test_fixbc_holds_its_two_states():
    """fixbc (31): face 3 at rho 2, p 1 and face 4 at rho 1, p 2.5, at rest,
    T = p/rho*const2; faces 1 and 2 left alone."""
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, 3, c, d)
    const2 = GAMMA * MACH * MACH
    I = slice(HM, HM + dims[0] + 1)
    for face, j, r, p in ((3, HM, 2.0, 1.0), (4, HM + dims[1], 1.0, 2.5)):
        o = box.fixbc(st, face)
        assert np.all(o["rho"][I, j, HM] == r) and np.all(o["prs"][I, j, HM] == p)
        assert np.all(o["vel"][I, j, HM, :] == 0.0)
        assert np.max(np.abs(o["tmp"][I, j, HM] - p / r * const2)) <= 1e-15
        assert np.max(np.abs(o["q"][I, j, HM, 4] - p / (GAMMA - 1.0))) <= 1e-15 * p
    for face in (1, 2):
        o = box.fixbc(st, face)
        for nm in ("q", "rho", "vel", "prs", "tmp"):
            assert np.array_equal(o[nm], st[nm])


@pytest.mark.parametrize("face", [3, 4])
def test_slipadib_wall_keeps_the_tangential_velocity(face):
    """slipadibwall (421): pressure, temperature and u by the zero-slope
    parabola's face value, w zero, v zero on face 3 and extrapolated alike on
    face 4, the density from the equation of state (not extrapolated) and
    the energy by the pressure form; faces 1 and 2 left alone."""
    dims = (12, 10, 0)
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, face, c, d)
    o = box.slipadib(st, face)
    hi = face == 4
    I = slice(HM, HM + dims[0] + 1)
    j0 = HM + (dims[1] if hi else 0)
    j1, j2 = (j0 - 1, j0 - 2) if hi else (j0 + 1, j0 + 2)
    const2 = GAMMA * MACH * MACH
    u = [o["vel"][I, j0, HM, m] for m in range(3)]
    p, rho, t = o["prs"][I, j0, HM], o["rho"][I, j0, HM], o["tmp"][I, j0, HM]
    assert np.max(np.abs(u[0] - c[1])) <= 1e-14
    if hi:
        assert np.max(np.abs(u[1] - c[2])) <= 1e-14
    else:
        assert np.all(u[1] == 0.0)
    assert np.all(u[2] == 0.0)
    assert np.max(np.abs(p - c[4])) <= 1e-14 * c[4]
    t1, t2 = st["tmp"][I, j1, HM], st["tmp"][I, j2, HM]
    assert np.max(np.abs(t - (4.0 * t1 - t2) / 3.0)) <= 1e-14
    assert np.max(np.abs(rho - p / t * const2)) <= 1e-14
    # the parabolic density is not what the wall takes
    assert np.min(np.abs(rho - c[0])) > 1e-6
    q = o["q"][I, j0, HM]
    assert np.array_equal(q[..., 0], rho)
    for m in range(3):
        assert np.array_equal(q[..., 1 + m], rho * u[m])
    ke = 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    assert np.max(np.abs(q[..., 4] - (p / (GAMMA - 1.0) + rho * ke))) <= 1e-14 * np.max(p)
    for f in (1, 2):
        o = box.slipadib(st, f)
        for nm in ("q", "rho", "vel", "prs", "tmp"):
            assert np.array_equal(o[nm], st[nm])


def test_slipadib_wall_only_on_faces_3_and_4():
    ob = po.OracleBox((16, 12, 10), variant=0, npdc=(4, 4, 3))
    ob.set_bc((0, 0, 421, 421, 0, 0))
    for f in (0, 1, 4, 5):
        bc = [0] * 6
        bc[f] = 421
        with pytest.raises(ValueError):
            ob.set_bc(tuple(bc))


def test_slipisot_wall_slips_up_to_xslip():
    """slipisotwall (411) on face 3: where the node's gridcube x = lx/im*i is
    at most xslip, the slip wall's node (u and T on the zero-slope parabola,
    v = w = 0); past it, at rest at the face's wall temperature; the pressure
    on the parabola and the density from the equation of state throughout.
    Face 4 is the isothermal noslip wall."""
    dims = (12, 10, 0)
    tw = (1.3, 1.1, 0.9, 1.2, 1.0, 1.0)
    xslip = 0.5  # nodes 0..6 slip, x(6) = 0.5 exactly
    box = po.UwdBox(dims, npdc=(4, 4, 3), gamma=GAMMA, mach=MACH, twall=tw, xslip=xslip)
    c = (1.1, 0.3, -0.2, 0.05, 1.0 / (GAMMA * MACH * MACH))
    d = (0.01, 0.02, -0.015, 0.005, 0.03)
    st, _ = parabolic_state(dims, 3, c, d)
    o = box.slipisot(st, 3)
    const2 = GAMMA * MACH * MACH
    I = slice(HM, HM + dims[0] + 1)
    slip = np.arange(dims[0] + 1) <= 6
    u = o["vel"][I, HM, HM, :]
    p, rho, t = o["prs"][I, HM, HM], o["rho"][I, HM, HM], o["tmp"][I, HM, HM]
    assert np.max(np.abs(p - c[4])) <= 1e-14 * c[4]
    assert np.max(np.abs(u[slip, 0] - c[1])) <= 1e-14
    assert np.all(u[~slip, 0] == 0.0) and np.all(u[:, 1:] == 0.0)
    t1, t2 = st["tmp"][I, HM + 1, HM], st["tmp"][I, HM + 2, HM]
    assert np.max(np.abs(t[slip] - (4.0 * t1[slip] - t2[slip]) / 3.0)) <= 1e-14
    assert np.all(t[~slip] == tw[2])
    assert np.max(np.abs(rho - p / t * const2)) <= 1e-14
    q = o["q"][I, HM, HM]
    assert np.array_equal(q[..., 1], rho * u[:, 0])
    assert np.max(np.abs(q[..., 4] - (p / (GAMMA - 1.0) + 0.5 * rho * u[:, 0] ** 2))) \
        <= 1e-14 * np.max(p)
    st4, _ = parabolic_state(dims, 4, c, d)
    a, b = box.slipisot(st4, 4), box.wall(st4, 4, True)
    for nm in ("q", "rho", "vel", "prs", "tmp"):
        assert np.array_equal(a[nm], b[nm])
    ob = po.OracleBox((16, 12, 10), variant=0, npdc=(4, 4, 3))
    ob.set_bc((0, 0, 411, 411, 0, 0))
    with pytest.raises(ValueError):
        ob.set_bc((411, 0, 0, 0, 0, 0))
