"""The checker's box with general metrics (all nine dxi components), pinned
on a periodic sheared grid: x = h(i + j), y = h j, z = h k.

Its nodes are those of the uniform grid, visited in another order (node
(i, j) of the sheared grid is node ((i + j) mod n, j) of the uniform one),
and its metrics are constant: d(xi)/dx = 1/h, d(xi)/dy = -1/h, d(eta)/dy =
1/h, d(zeta)/dz = 1/h, J = h^3.  The same physical field sampled on both
must give
  - gradients that match the analytic ones (a transposed or misplaced
    metric component shows at O(1));
  - a right-hand side that, node for node, matches the uniform grid's up
    to the schemes' truncation error (the lines run along other
    directions)."""
import math

import numpy as np
import pytest

import pyoracle as po

HM = 5
N = 48


def sheared_metrics(n, s=1.0):
    h = 2 * math.pi / n
    shape = (n + 1 + 2 * HM,) * 3
    jac = np.full(shape, h * h * h)
    dxi = np.zeros(shape + (3, 3))
    dxi[..., 0, 0] = 1.0 / h
    dxi[..., 0, 1] = -s / h
    dxi[..., 1, 1] = 1.0 / h
    dxi[..., 2, 2] = 1.0 / h
    return jac, dxi


def tgv_fields(X, Y, Z, mach=0.1, gamma=1.4):
    pinf = 1.0 / (gamma * mach * mach)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = 0.0 * X
    p = pinf + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    rho = 1.0 + 0.0 * X
    return rho, u, v, w, p


def load(ob, X, Y, Z):
    rho, u, v, w, p = tgv_fields(X, Y, Z)
    st = po.prim_to_state(rho, u, v, w, p, mach=0.1)
    for m in range(5):
        ob.set(f"q{m + 1}", st["q"][..., m])
    ob.set("rho", rho)
    for nm, a in zip("uvw", (u, v, w)):
        ob.set(nm, a)
    ob.set("prs", p)
    ob.set("tmp", st["tmp"])


def nodes(n, sheared):
    h = 2 * math.pi / n
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    return h * (i + j if sheared else i), h * j, h * k


def test_gradients_on_a_sheared_grid_match_the_analytic_ones():
    ob = po.OracleBox(N, variant=0)
    ob.set_metrics(*sheared_metrics(N))
    X, Y, Z = nodes(N, True)
    load(ob, X, Y, Z)
    ob.qswap()
    ob.rhscal()
    exact = {
        (0, 0): np.cos(X) * np.cos(Y) * np.cos(Z), (0, 1): -np.sin(X) * np.sin(Y) * np.cos(Z),
        (0, 2): -np.sin(X) * np.cos(Y) * np.sin(Z), (1, 0): np.sin(X) * np.sin(Y) * np.cos(Z),
        (1, 1): -np.cos(X) * np.cos(Y) * np.cos(Z), (1, 2): np.cos(X) * np.sin(Y) * np.sin(Z),
    }
    for (m, e), ref in exact.items():
        d = ob.deriv(3 * m + e)
        assert np.max(np.abs(d - ref)) < 1e-5, (m, e, np.max(np.abs(d - ref)))


def test_rhs_on_a_sheared_grid_matches_the_uniform_grid():
    a = po.OracleBox(N, variant=0)
    a.set_metrics(*sheared_metrics(N))
    load(a, *nodes(N, True))
    b = po.OracleBox(N, variant=0)
    load(b, *nodes(N, False))
    for ob in (a, b):
        ob.qswap()
        ob.rhscal()
    i = np.arange(N + 1)[:, None, None]
    j = np.arange(N + 1)[None, :, None]
    ip = (i + j) % N
    # the residual is a small difference of fluxes of size J p d(xi)/dx
    # (p = 1/(gamma Ma^2) = 71): a misplaced metric term shows at that
    # scale, truncation error far below it
    h = 2 * math.pi / N
    scale = h ** 3 * float(np.max(np.abs(a.get("prs")))) / h
    for m in range(5):
        qa = a.deriv(12 + m)
        qb = b.deriv(12 + m)[ip, j, np.arange(N + 1)[None, None, :]]
        assert np.max(np.abs(qa - qb)) < 1e-6 * scale, (m, np.max(np.abs(qa - qb)) / scale)
