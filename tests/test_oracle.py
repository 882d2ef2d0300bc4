"""Pins the C restatement (oracle/oracle.c) before it is trusted as the checker.

Each part is checked against something that would fail on a plausible
mistake (dropped term, wrong sign or index, transposed operand):
  * the reference's own modules compiled from source (golden vectors in
    tests/golden/lines_ref.npz, and the live library where it is built),
  * the reference-held series miniapps/tgv_solver_3d/state.ref_128,
  * mathematics: dense solves, polynomial exactness, transfer functions,
    free-stream preservation, closed-form initial statistics.
"""
import math
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "lines_ref.npz"))
HM = 5


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("dim", [64, 128, 256])
@pytest.mark.parametrize("nt", [1, 2, 3, 4])
@pytest.mark.parametrize("tag", ["rand", "smooth"])
def test_line_ops_match_reference_golden(dim, nt, tag):
    k = f"{tag}_{nt}_{dim}"
    f = G["f_" + k]
    assert np.array_equal(pyoracle.df_compact(nt, f), G["df_" + k])
    assert np.array_equal(pyoracle.diff6ec(nt, f), G["d6_" + k])
    assert np.array_equal(pyoracle.compact_filter(nt, f, 0.49, 1.11, 0.98), G["ff098_" + k])
    assert np.array_equal(pyoracle.compact_filter(nt, f, 0.49, 1.11, 1.09), G["ff109_" + k])


@pytest.mark.parametrize("dim", [64, 128, 256])
@pytest.mark.parametrize("nt", [1, 2, 3, 4])
def test_lhs_tables_match_reference(dim, nt):
    for op, key in (("deriv", "fdac"), ("filter", "flac")):
        _, _, _, _, ac = pyoracle.line_system(op, nt, dim)
        assert np.array_equal(ac, G[f"{key}_{nt}_{dim}"])


def test_filter_coefficients_match_reference():
    assert np.array_equal(pyoracle.filter_coef(0.49, 1.11, 0.98), G["coef_098"])
    assert np.array_equal(pyoracle.filter_coef(0.49, 1.11, 1.09), G["coef_109"])


@pytest.mark.parametrize("n", [3, 5, 8, 16])
def test_thomas_matches_reference_and_dense_solve(n):
    a, c, d = G[f"thomas_a_{n}"], G[f"thomas_c_{n}"], G[f"thomas_d_{n}"]
    x = pyoracle.thomas(a, c, d)
    assert np.array_equal(x, G[f"thomas_x_{n}"])
    # dense matrix of the same system: unit diagonal, a below, c above, and
    # the first row's a ignored (src/commfunc.F90:760)
    A = np.eye(n)
    for i in range(n):
        if i > 0:
            A[i, i - 1] = a[i]
        if i < n - 1:
            A[i, i + 1] = c[i]
    assert np.allclose(A @ x, d, rtol=0, atol=1e-13)


def test_live_reference_library_if_built():
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    rng = np.random.default_rng(7)
    for dim in (16, 33, 100):
        for nt in (1, 2, 3, 4):
            f = rng.standard_normal(dim + 11)
            assert np.array_equal(pyoracle.df_compact(nt, f), pyoracle.df_compact(nt, f, lib=ref))
            assert np.array_equal(pyoracle.compact_filter(nt, f),
                                  pyoracle.compact_filter(nt, f, lib=ref))


# ------------------------------------------------------------ mathematics

def _poly_line(dim, deg, h=1.0):
    x = np.arange(-HM, dim + HM + 1) * h
    return x, x ** deg


@pytest.mark.parametrize("nt", [1, 2, 3, 4])
def test_derivative_exact_on_polynomials(nt):
    # interior 6th-order compact and explicit interface rows are exact up to
    # degree 6 in index space; the 3rd-order boundary rows up to degree 3
    dim = 40
    for deg in range(0, 4):
        x, f = _poly_line(dim, deg)
        df = pyoracle.df_compact(nt, f)
        exact = deg * x[HM:HM + dim + 1] ** max(deg - 1, 0) if deg else np.zeros(dim + 1)
        assert np.allclose(df, exact, rtol=1e-11, atol=1e-9), deg
    if nt == 3:
        for deg in (4, 5, 6):
            x, f = _poly_line(dim, deg, h=0.05)
            df = pyoracle.df_compact(nt, f)
            exact = deg * x[HM:HM + dim + 1] ** (deg - 1) * 0.05
            assert np.allclose(df, exact, rtol=1e-10, atol=1e-12), deg


def test_derivative_modified_wavenumber():
    # periodic data: away from the closure rows the compact scheme returns
    # k'h sin/cos with k'h = (14/9 sin t + 1/18 sin 2t) / (1 + 2/3 cos t)
    dim = 64
    for kw in (1, 3, 8):
        th = 2 * np.pi * kw / dim
        nodes = np.arange(-HM, dim + HM + 1)
        f = np.sin(th * nodes)
        df = pyoracle.df_compact(3, f)
        kh = (14 / 9 * np.sin(th) + 1 / 18 * np.sin(2 * th)) / (1 + 2 / 3 * np.cos(th))
        mid = slice(24, 41)
        assert np.allclose(df[mid], kh * np.cos(th * np.arange(dim + 1))[mid], atol=1e-10)


def test_derivative_antisymmetry():
    rng = np.random.default_rng(3)
    f = rng.standard_normal(75)
    assert np.array_equal(pyoracle.df_compact(3, -f), -pyoracle.df_compact(3, f))


def test_filter_transfer_function_and_constants():
    dim = 64
    nodes = np.arange(-HM, dim + HM + 1)
    # constants pass unchanged (T(0) = 1) through every closure
    for nt in (1, 2, 3, 4):
        ff = pyoracle.compact_filter(nt, np.full(dim + 11, 2.5))
        assert np.allclose(ff, 2.5, rtol=0, atol=1e-13)
    # the odd-even mode is removed in the interior (T(pi) = 0)
    ff = pyoracle.compact_filter(3, (-1.0) ** nodes)
    assert np.max(np.abs(ff[20:45])) < 1e-12
    # a low mode is transmitted with T(t) = sum a_n cos(n t) / (1 + 2 alpha cos t)
    al = 0.49
    a = [(193 + 126 * al) / 256, (105 + 302 * al) / 256, (-15 + 30 * al) / 64,
         (45 - 90 * al) / 512, (-5 + 10 * al) / 256, (1 - 2 * al) / 512]
    th = 2 * np.pi * 5 / dim
    T = sum(a[n] * np.cos(n * th) for n in range(6)) / (1 + 2 * al * np.cos(th))
    ff = pyoracle.compact_filter(3, np.cos(th * nodes))
    assert np.allclose(ff[20:45], T * np.cos(th * np.arange(dim + 1))[20:45], atol=1e-13)


def test_explicit_derivative_exact_on_polynomials():
    # 2nd-order one-sided/central boundary rows are exact on quadratics, the
    # 6th-order interior up to degree 6
    x, f = _poly_line(30, 2)
    for nt in (1, 2, 3, 4):
        d = pyoracle.diff6ec(nt, f)
        assert np.allclose(d, 2 * x[HM:HM + 31], rtol=1e-12, atol=1e-9)
    for deg in (3, 6):
        x, f = _poly_line(30, deg, h=0.1)
        d = pyoracle.diff6ec(3, f)
        assert np.allclose(d, deg * x[HM:HM + 31] ** (deg - 1) * 0.1, rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------ whole solver

def test_miniapp_variant_reproduces_state_ref_128_first_steps():
    """The reference-held series (128^3, dt 1e-3): first steps, bit for bit."""
    ref = np.loadtxt(os.path.join(GOLDEN, "state.ref_128.txt"))
    b = pyoracle.OracleBox(128, variant=1)
    b.init_tgv()
    for s in range(3):
        ke, en, _ = b.step(1e-3, stats=True)
        assert ke == ref[s, 2] and en == ref[s, 3], (s, ke, en)
    b.close()


@pytest.mark.slow
def test_miniapp_variant_reproduces_state_ref_128_all_steps():
    ref = np.loadtxt(os.path.join(GOLDEN, "state.ref_128.txt"))
    b = pyoracle.OracleBox(128, variant=1)
    b.init_tgv()
    for s in range(ref.shape[0]):
        ke, en, _ = b.step(1e-3, stats=True)
        assert ke == ref[s, 2] and en == ref[s, 3], (s, ke, en)


def test_bridge_variant_reproduces_state_ref_128_first_steps():
    """The main path's right-hand side and RK3 combine (gradcal's metric
    products, the J-weighted convective and viscous fluxes in metric form,
    qsave = q J and the combine divided by J) driven in the mini-app's
    stage order: the reference-held series within 1e-12 relative.  A
    dropped J, a wrong metric component or a missing pressure term in the
    main path's flux assembly moves the series by far more."""
    ref = np.loadtxt(os.path.join(GOLDEN, "state.ref_128.txt"))
    b = pyoracle.OracleBox(128, variant=2)
    b.init_tgv()
    for s in range(3):
        ke, en, _ = b.step(1e-3, stats=True)
        assert abs(ke - ref[s, 2]) <= 1e-12 * ref[s, 2], (s, ke, ref[s, 2])
        assert abs(en - ref[s, 3]) <= 1e-12 * ref[s, 3], (s, en, ref[s, 3])
    b.close()


@pytest.mark.slow
def test_bridge_variant_reproduces_state_ref_128_all_steps():
    ref = np.loadtxt(os.path.join(GOLDEN, "state.ref_128.txt"))
    b = pyoracle.OracleBox(128, variant=2)
    b.init_tgv()
    for s in range(ref.shape[0]):
        ke, en, _ = b.step(1e-3, stats=True)
        assert abs(ke - ref[s, 2]) <= 1e-12 * ref[s, 2], (s, ke, ref[s, 2])
        assert abs(en - ref[s, 3]) <= 1e-12 * ref[s, 3], (s, en, ref[s, 3])


def test_main_variant_initial_statistics_closed_form():
    # KE0 = 1/8, enstrophy0 = 3/8 for the Taylor-Green field (the first row of
    # state.ref_128 holds the same values); dissipation0 = 2 nu * 3/8 up to
    # compressibility at Ma 0.1
    b = pyoracle.OracleBox(32, variant=0)
    b.init_tgv()
    b.filterq()
    b.qswap()
    b.rhscal()
    ke, en, di = b.stats()
    assert abs(ke - 0.125) < 1e-9
    assert abs(en - 0.375) / 0.375 < 2e-3
    assert abs(di - 2 * 0.375 / 1600) / (2 * 0.375 / 1600) < 2e-2


def test_statistics_closed_forms_on_a_compressible_field():
    """kenergycal, enstophycal and diss_rate_cal (src/statistic.F90:870-1042)
    on u = a sin x + d sin y, v = b sin y, w = c sin z, rho = T = 1, a field
    with div u = a cos x + b cos y + c cos z != 0 and one shear component.
    With mu = 1/Re at T = 1 the means over the periodic nodes are
      KE = (a^2 + b^2 + c^2 + d^2) / 4,   enstrophy = d^2 / 4,
      dissipation = 2 mu <S:S - (div u)^2 / 3> = (2/3) mu (a^2+b^2+c^2) + mu d^2 / 2,
    since <S:S> = (a^2+b^2+c^2)/2 + d^2/4 and <(div u)^2> = (a^2+b^2+c^2)/2.
    Dropping the dilatation term gives mu (a^2+b^2+c^2) + mu d^2/2, off by
    a fifth; the gradients are 6th-order compact derivatives of single
    modes at 64 points (error ~1e-9), so the bar is 1e-8."""
    n, a, bb, c, d = 64, 0.3, -0.2, 0.25, 0.4
    box = pyoracle.OracleBox(n, variant=0)
    x = 2 * np.pi / n * np.arange(n + 1)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = a * np.sin(X) + d * np.sin(Y)
    v = bb * np.sin(Y)
    w = c * np.sin(Z)
    gamma, mach, re = 1.4, 0.1, 1600.0
    const1 = 1.0 / (gamma * (gamma - 1.0) * mach * mach)
    one = np.ones_like(u)
    for nm, f in (("q1", one), ("q2", u), ("q3", v), ("q4", w),
                  ("q5", const1 + 0.5 * (u * u + v * v + w * w)), ("rho", one), ("u", u),
                  ("v", v), ("w", w), ("prs", one / (gamma * mach * mach)), ("tmp", one)):
        box.set(nm, f)
    box.qswap()
    box.rhscal()
    ke, en, di = box.stats()
    box.close()
    mu = 1.0 / re
    s2 = a * a + bb * bb + c * c
    assert abs(ke - (s2 + d * d) / 4) <= 1e-13 * ke
    assert abs(en - d * d / 4) <= 1e-8 * en, en
    want = 2.0 / 3.0 * mu * s2 + 0.5 * mu * d * d
    assert abs(di - want) <= 1e-8 * want, (di, want)
    assert abs(di - (mu * s2 + 0.5 * mu * d * d)) > 0.1 * want


def test_main_variant_free_stream_preservation():
    # a uniform state has zero right-hand side through every operator
    n = 24
    b = pyoracle.OracleBox(n, variant=0)
    b.init_tgv()
    shape = b.shape
    const1 = 1 / (1.4 * 0.4 * 0.01)
    b.set("q1", np.full(shape, 1.2))
    b.set("q2", np.full(shape, 1.2 * 0.3))
    b.set("q3", np.full(shape, -1.2 * 0.1))
    b.set("q4", np.full(shape, 1.2 * 0.2))
    T = 1.05
    b.set("q5", np.full(shape, 1.2 * (T * const1 + 0.5 * (0.09 + 0.01 + 0.04))))
    b.lib.orc_box_updatefvar(b.h)
    b.filterq()
    b.qswap()
    b.rhscal()
    for m in range(5):
        assert np.max(np.abs(b.deriv(12 + m))) < 1e-9
    # the filter keeps the uniform state
    assert np.allclose(b.get("q1"), 1.2, rtol=0, atol=1e-13)


def test_qswap_semantics():
    # after qswap the duplicate nodes agree and the primitives on the seam
    # planes are those of the averaged q; interior primitives are left stale
    n = 20
    b = pyoracle.OracleBox(n, variant=0)
    b.init_tgv()
    b.filterq()
    stale_u = b.get("u")
    b.qswap()
    q2, q1, u = b.get("q2"), b.get("q1"), b.get("u")
    assert np.array_equal(q2[0], q2[-1]) and np.array_equal(q2[:, 0], q2[:, -1])
    assert np.array_equal(q2[:, :, 0], q2[:, :, -1])
    assert np.array_equal(u[0], q2[0] / q1[0])
    assert np.array_equal(u[5:-5, 5:-5, 5:-5], stale_u[5:-5, 5:-5, 5:-5])
    assert not np.array_equal(u[5:-5, 5:-5, 5:-5], (q2 / q1)[5:-5, 5:-5, 5:-5])
