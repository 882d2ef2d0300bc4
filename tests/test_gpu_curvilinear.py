"""The central path on a grid with general metrics (all nine dxi
components varying from node to node) against the checker, which reads the
same uploaded metrics: gradients and right-hand side of one stage, then
whole RK3 steps.

The grid is a periodic wavy one, x = h i + eps sin(2 pi j/nj), y = h j +
eps sin(2 pi k/nk), z = h k + eps sin(2 pi i/ni); its metrics come from the
analytic inverse of d(x,y,z)/d(i,j,k) at every node, J its determinant.
The checker's general-metric forms are pinned separately on a sheared
grid (tests/test_oracle_curvilinear.py)."""
import math
import os
import sys

import numpy as np
import pytest

import astr_gpu_amd as ag

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import pyoracle  # noqa: E402

pytestmark = pytest.mark.gpu
HM = 5


def wavy_grid(dims, eps=0.05):
    n = np.array(dims, dtype=float)
    h = 2 * math.pi / n
    idx = [np.arange(-HM, d + HM + 1, dtype=float) for d in dims]
    I, J, K = np.meshgrid(*idx, indexing="ij")
    w = 2 * math.pi / n
    X = h[0] * I + eps * np.sin(w[1] * J)
    Y = h[1] * J + eps * np.sin(w[2] * K)
    Z = h[2] * K + eps * np.sin(w[0] * I)
    A = np.zeros(I.shape + (3, 3))  # A[..., x_e, xi_d] = d x_e / d xi_d
    A[..., 0, 0] = h[0]
    A[..., 0, 1] = eps * w[1] * np.cos(w[1] * J)
    A[..., 1, 1] = h[1]
    A[..., 1, 2] = eps * w[2] * np.cos(w[2] * K)
    A[..., 2, 2] = h[2]
    A[..., 2, 0] = eps * w[0] * np.cos(w[0] * I)
    jac = np.linalg.det(A)
    dxi = np.linalg.inv(A)  # dxi[..., d, e] = d xi_d / d x_e
    return X, Y, Z, np.asfortranarray(jac), np.asfortranarray(dxi)


def state(X, Y, Z, mach=0.1):
    pinf = 1.0 / (1.4 * mach * mach)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = 0.1 * np.sin(X + Z)
    p = pinf + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    return pyoracle.prim_to_state(1.0 + 0.0 * X, u, v, w, p, mach=mach)


def pair(dims):
    X, Y, Z, jac, dxi = wavy_grid(dims)
    st = state(X, Y, Z)
    I = tuple(slice(HM, HM + d + 1) for d in dims)
    ob = pyoracle.OracleBox(dims, variant=0)
    ob.set_metrics(jac, dxi)
    for m in range(5):
        ob.set(f"q{m + 1}", st["q"][I + (m,)])
    ob.set("rho", st["rho"][I])
    for c, nm in enumerate("uvw"):
        ob.set(nm, st["vel"][I + (c,)])
    ob.set("prs", st["prs"][I])
    ob.set("tmp", st["tmp"][I])
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    s.upload_geom(jac, dxi)
    z = {k: np.array(v) for k, v in st.items()}
    for k in z:  # halos as the checker starts them: zero until qswap
        a = np.zeros_like(z[k])
        a[I] = z[k][I]
        z[k] = np.asfortranarray(a)
    s.upload_state(z["q"], z["rho"], z["vel"], z["prs"], z["tmp"])
    return ob, s, p


def nerr(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def test_curvilinear_stage_matches_checker():
    dims = (40, 36, 32)
    print(f"case curvilinear stage dims={dims}", flush=True)
    ob, s, p = pair(dims)
    ob.qswap()
    s.qswap()
    ob.rhscal()
    s.rhscal()
    r = s.download_rhs()
    # each gradient component against the scale of the whole gradient (a
    # component that vanishes in the continuum, as dw/dy here, is a
    # cancellation of the axes' terms and no scale of its own)
    ref = {(m, e): ob.deriv(3 * m + e) for m in range(3) for e in range(3)}
    gscale = max(float(np.max(np.abs(v))) for v in ref.values())
    worst = 0.0
    for (m, e), v in ref.items():
        ed = float(np.max(np.abs(r["dvel"][..., m, e] - v))) / gscale
        worst = max(worst, ed)
        assert ed < 1e-12, ("dvel", m, e, ed)
    tref = [ob.deriv(9 + e) for e in range(3)]
    tscale = max(float(np.max(np.abs(v))) for v in tref)
    for e in range(3):
        assert float(np.max(np.abs(r["dtmp"][..., e] - tref[e]))) / tscale < 1e-12, ("dtmp", e)
    # the residual is a small difference of fluxes of size J p d(xi)/dx
    flux_scale = float(np.max(np.abs(ob.get("prs")))) * (2 * math.pi / min(dims)) ** 2
    for m in range(5):
        e = float(np.max(np.abs(r["qrhs"][..., m] - ob.deriv(12 + m)))) / flux_scale
        print(f"  qrhs{m + 1}: {e:.2e} of the flux scale", flush=True)
        assert e < 1e-12, ("qrhs", m, e)
    print(f"  gradients {worst:.2e}", flush=True)
    s.close()


def test_curvilinear_steps_match_checker():
    dims = (40, 36, 32)
    print(f"case curvilinear steps dims={dims}", flush=True)
    ob, s, p = pair(dims)
    dt = 0.004
    for n in range(3):
        so = ob.step(dt, stats=True)
        sg = s.time_integration_rk(dt, stats=True)
        for a, b in zip(sg, so):
            assert abs(a - b) <= 1e-10 * abs(b), (n, sg, so)
    out = s.download_state()
    I = tuple(slice(HM, HM + d + 1) for d in dims)
    ref = [ob.get(f"q{m + 1}") for m in range(5)]
    mom = max(float(np.max(np.abs(ref[m]))) for m in (1, 2, 3))
    for m in range(5):
        scale = mom if m in (1, 2, 3) else float(np.max(np.abs(ref[m])))
        e = float(np.max(np.abs(out["q"][I + (m,)] - ref[m]))) / scale
        print(f"  q{m + 1}: {e:.2e}", flush=True)
        assert e < 1e-10, (m, e)
    s.close()
