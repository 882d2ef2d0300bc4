"""The checker's box with bounded axes (npdc 4: a single rank's axis with a
boundary at both ends).

Its line operators are the ones pinned against the reference's modules
(tests/test_oracle.py); what these tests pin is the composition: on a
bounded axis qswap and the halo copies do nothing, and the boundary
closures keep a uniform flow uniform through whole steps."""
import numpy as np
import pytest

import pyoracle

DIMS = (24, 20, 18)


def uniform(ob, u=(0.3, -0.2, 0.1), rho=1.2, p=80.0, gamma=1.4, mach=0.1):
    n = ob.shape
    const1 = 1.0 / (gamma * (gamma - 1.0) * mach * mach)
    const2 = gamma * mach * mach
    t = p / rho * const2
    ob.set("rho", np.full(n, rho))
    for c, nm in enumerate("uvw"):
        ob.set(nm, np.full(n, u[c]))
    ob.set("prs", np.full(n, p))
    ob.set("tmp", np.full(n, t))
    q = [rho, rho * u[0], rho * u[1], rho * u[2], rho * (t * const1 + 0.5 * sum(v * v for v in u))]
    for m in range(5):
        ob.set(f"q{m + 1}", np.full(n, q[m]))
    return q


@pytest.mark.parametrize("npdc", [(4, 4, 4), (4, 3, 4), (3, 3, 4)])
def test_uniform_flow_stays_uniform(npdc):
    ob = pyoracle.OracleBox(DIMS, variant=0, npdc=npdc)
    q = uniform(ob)
    for _ in range(2):
        ob.step(0.002)
    for m in range(5):
        a = ob.get(f"q{m + 1}")
        assert np.max(np.abs(a - q[m])) <= 1e-12 * abs(q[m]) + 1e-14, (npdc, m)


def test_qswap_leaves_a_bounded_axis_alone():
    ob = pyoracle.OracleBox(DIMS, variant=0, npdc=(4, 3, 4))
    ob.init_tgv()
    rng = np.random.default_rng(5)
    a = ob.get("q2") + 1e-3 * rng.standard_normal(ob.shape)
    ob.set("q2", a)
    before = ob.get_haloed("q2")
    ob.qswap()
    after = ob.get_haloed("q2")
    H = pyoracle.HM
    n0, n1, n2 = DIMS
    # i and k bounded (npdc 4): their halos and end nodes untouched
    # away from the j seam; j periodic: its duplicates averaged
    inner_j = slice(H + 1, H + n1)
    assert np.array_equal(after[:H, inner_j, H:H + n2 + 1], before[:H, inner_j, H:H + n2 + 1])
    assert np.array_equal(after[H + n0, inner_j, H + 1:H + n2], before[H + n0, inner_j, H + 1:H + n2])
    assert np.array_equal(after[H:H + n0 + 1, inner_j, :H], before[H:H + n0 + 1, inner_j, :H])
    j0, jn = after[:, H, :], after[:, H + n1, :]
    assert np.array_equal(j0[H:H + n0 + 1, H:H + n2 + 1], jn[H:H + n0 + 1, H:H + n2 + 1])
    assert not np.array_equal(before[:, H, :], before[:, H + n1, :])


def test_bounded_closures_need_the_main_variant():
    with pytest.raises(ValueError):
        pyoracle.OracleBox(DIMS, variant=1, npdc=(4, 3, 3))


def pulse(ob, mach, u=0.3, v=0.2, amp=0.05, gamma=1.4):
    """Uniform flow at the freestream pressure 1/(gamma Ma^2) with a smooth
    pressure pulse near the x-high and y-high faces; the sound speed
    sqrt(T)/Ma stays near 2.5 at Ma 0.4, away from the integer values
    gcnscbc truncates it to."""
    n = ob.shape
    x, y, z = np.meshgrid(*[np.linspace(0.0, 1.0, k) for k in n], indexing="ij")
    const1 = 1.0 / (gamma * (gamma - 1.0) * mach * mach)
    const2 = gamma * mach * mach
    g = amp * np.exp(-((x - 0.8) ** 2 + (y - 0.7) ** 2 + (z - 0.5) ** 2) / 0.03)
    rho = 1.0 + g
    vel = (u + g, v - 0.5 * g, 0.1 * g)
    p = (1.0 + g) / const2
    t = p / rho * const2
    ob.set("rho", rho)
    for c, nm in enumerate("uvw"):
        ob.set(nm, vel[c])
    ob.set("prs", p)
    ob.set("tmp", t)
    q = [rho, rho * vel[0], rho * vel[1], rho * vel[2],
         rho * (t * const1 + 0.5 * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))]
    for m in range(5):
        ob.set(f"q{m + 1}", q[m])


GCN_FACES = (0, 23, 0, 23, 0, 0)


def test_gcnscbc_keeps_the_freestream():
    """gcnscbc (bctype 23) on the x-high and y-high faces of a three-dimensional
    box (bounded x and y, periodic z) through whole central steps: at the
    freestream pressure with T = 1 every LODI term is zero and the boundary
    node takes its neighbour's state, so the flow stays uniform."""
    mach = 0.4
    ob = pyoracle.OracleBox(DIMS, variant=0, npdc=(4, 4, 3), mach=mach)
    ob.set_bc(GCN_FACES)
    q = uniform(ob, u=(0.3, 0.2, 0.1), rho=1.0, p=1.0 / (1.4 * mach * mach), mach=mach)
    for _ in range(2):
        ob.step(0.01)
    for m in range(5):
        a = ob.get(f"q{m + 1}")
        assert np.max(np.abs(a - q[m])) <= 1e-12 * abs(q[m]) + 1e-14, m


def test_gcnscbc_sets_the_boundary_nodes():
    """With a pulse reaching the faces, a step with gcnscbc differs from the
    plain bounded step most on the boundary planes it rewrites, and the
    difference fades away from them (a zero time step: what moves is the
    boundary planes and, through the implicit filter of the later stages,
    the lines that end on them)."""
    mach = 0.4
    runs = {}
    for faces in ((0,) * 6, GCN_FACES):
        ob = pyoracle.OracleBox(DIMS, variant=0, npdc=(4, 4, 3), mach=mach)
        ob.set_bc(faces)
        pulse(ob, mach)
        ob.step(0.0)
        runs[faces[1]] = {nm: ob.get(nm) for nm in ("q1", "q2", "prs", "tmp")}
    plain, gcn = runs[0], runs[23]
    for nm in ("q1", "q2", "prs", "tmp"):
        d = np.abs(gcn[nm] - plain[nm])
        face = max(np.max(d[-1, :, :]), np.max(d[:, -1, :]))
        assert face > 1e-6, nm
        assert np.max(d[: DIMS[0] // 2, : DIMS[1] // 2, :]) < 1e-2 * face, nm


def test_set_bc_rejects_unbuilt_types():
    ob = pyoracle.OracleBox(DIMS, variant=0, npdc=(4, 4, 3))
    with pytest.raises(ValueError):
        ob.set_bc((0, 22, 0, 0, 0, 0))
