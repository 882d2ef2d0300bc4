"""The adjoint building blocks on the GPU (through the C ABI).

The transposed line operators are checked two ways: the dot-product
identity <D f, g> = <f, D^T g> on random fields of every closure type, and
node for node against the explicit transpose of the checker's line
operators (which are bit-identical to the reference modules), built by
applying them to unit vectors.
"""
import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle as po

pytestmark = pytest.mark.gpu
HM = 5


def forward(s, op, d, f):
    return s.central(d, f) if op == 0 else s.compact_filter(d, f)


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("ntype", [1, 2, 3, 4])
def test_line_adjoint_dot_product(op, ntype):
    dims = (40, 36, 33)
    print(f"case line adjoint op={op} ntype={ntype}", flush=True)
    p = ag.default_params(dims, npdci=ntype, npdcj=ntype, npdck=ntype)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(10 * op + ntype)
    shape = ag.haloed_shape(p)
    for d in (1, 2, 3):
        f = np.asfortranarray(rng.standard_normal(shape))
        g = np.asfortranarray(rng.standard_normal(shape))
        y = forward(s, op, d, f)
        fb = s.line_adjoint(op, d, g)
        lhs = float(np.sum(y * ag.interior(g, p)))
        rhs = float(np.sum(f * fb))
        scale = float(np.sum(np.abs(y * ag.interior(g, p))))
        print(f"  dir {d}: <Df,g> {lhs:.15e} <f,D'g> {rhs:.15e}")
        assert abs(lhs - rhs) <= 1e-13 * scale, (d, lhs, rhs)
        # only the halos along d (and the interior) can be nonzero
        other = [ax for ax in range(3) if ax != d - 1]
        for ax in other:
            sl = [slice(None)] * 3
            sl[ax] = slice(0, HM)
            assert not np.any(fb[tuple(sl)])
    s.close()


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("ntype", [1, 3, 4])
def test_line_adjoint_matches_the_checker_transpose(op, ntype):
    """Columns of the checker's operator M (df_compact or compact_filter on
    unit vectors, lines -hm..n+hm -> 0..n): the GPU transpose of g equals
    M^T g along every line."""
    n = 40
    dims = (n, 36, 33)
    p = ag.default_params(dims, npdci=ntype, npdcj=3, npdck=3)
    s = ag.AstrSolver(p)
    width = n + 1 + 2 * HM
    M = np.zeros((n + 1, width))
    for c in range(width):
        e = np.zeros(width)
        e[c] = 1.0
        M[:, c] = po.df_compact(ntype, e) if op == 0 else po.compact_filter(ntype, e)
    rng = np.random.default_rng(7)
    g = np.asfortranarray(rng.standard_normal(ag.haloed_shape(p)))
    fb = s.line_adjoint(op, 1, g)
    gi = ag.interior(g, p)
    ref = np.einsum("ic,ijk->cjk", M, gi)
    got = fb[:, HM:-HM, HM:-HM]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"case checker transpose op={op} ntype={ntype}: {err:.2e}")
    assert err <= 1e-13
    s.close()


@pytest.mark.parametrize("d", [1, 2, 3])
def test_filter_adjoint_with_periodic_halos(d):
    """wrap: the forward reads its halos as periodic copies of the nodes
    (filterq's dataswap); the transpose folds them back."""
    dims = (32, 36, 40)
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    rng = np.random.default_rng(d)
    shape = ag.haloed_shape(p)
    f = np.asfortranarray(rng.standard_normal(shape))
    # halos along d: f(-k) = f(n-k), f(n+k) = f(k)
    n = dims[d - 1]
    idx = np.arange(-HM, n + HM + 1)
    src = np.where(idx < 0, idx + n, np.where(idx > n, idx - n, idx)) + HM
    f = np.asfortranarray(np.take(f, src, axis=d - 1))
    g = np.asfortranarray(rng.standard_normal(shape))
    y = s.compact_filter(d, f)
    fb = s.line_adjoint(1, d, g, wrap=True)
    lhs = float(np.sum(y * ag.interior(g, p)))
    rhs = float(np.sum(ag.interior(f, p) * ag.interior(fb, p)))
    scale = float(np.sum(np.abs(y * ag.interior(g, p))))
    print(f"case wrap dir {d}: {lhs:.15e} {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-13 * scale
    assert not np.any(fb[:HM]) and not np.any(fb[-HM:])
    s.close()
