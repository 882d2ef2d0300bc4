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


# ---------------------------------------------------------------- full adjoint

def adjoint_case(dims, seed=3, lengths=(2 * np.pi,) * 3):
    import torch
    from torch_forward import Box, to_host

    p = ag.default_params(dims, mach=0.3, reynolds=173.0)
    st = ag.hit_state(p, seed)
    u0 = np.moveaxis(ag.interior(st["vel"], p)[:-1, :-1, :-1], -1, 0).copy()
    box = Box(dims, reynolds=p.reynolds, mach=p.mach, lengths=lengths)
    return p, box, torch.tensor(u0), to_host


@pytest.mark.parametrize("nsteps,ckpt,lengths,dims", [
    (2, 1, None, (16, 18, 20)), (2, 8, None, (16, 18, 20)), (3, 2, None, (16, 18, 20)),
    (2, 2, (2 * np.pi, 4.0, 9.0), (16, 18, 20)), (2, 2, None, (29, 16, 27))])
def test_adjoint_matches_torch_autograd(nsteps, ckpt, lengths, dims):
    """dJ/d(state) of the GPU reversal against PyTorch-CPU autograd of the
    same forward (RK3 steps, filter and viscous terms on): every stage kept,
    one segment of stored stages, a last segment shorter than ckpt, a box
    whose three axes have different spacings (metrics and J), and one whose
    x halo strip n+1..n+5 runs past the 128-byte line of node n (n = 29,
    n mod 16 = 13: the flux derivative reads G_x at n+3 and n+4 there)."""
    import torch

    lengths = lengths or (2 * np.pi,) * 3
    print(f"case adjoint vs autograd dims={dims} nsteps={nsteps} ckpt={ckpt} lengths={lengths}",
          flush=True)
    p, box, u0, to_host = adjoint_case(dims, lengths=lengths)
    dt = 0.004
    Q0, P0 = box.initial(u0)
    Q0 = Q0.detach().requires_grad_(True)
    P0 = P0.detach().requires_grad_(True)
    J = box.objective_state(Q0, P0, nsteps, dt)
    J.backward()
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p, lengths))
    h = to_host(Q0, P0)
    s.upload_state(h["q"], h["rho"], h["vel"], h["prs"], h["tmp"])
    Jg, qb, pb, ms = s.adjoint_enstrophy(nsteps, dt, ckpt)
    print(f"  J gpu {Jg:.15e} torch {float(J.detach()):.15e}; ms {ms}")
    assert abs(Jg - float(J)) <= 1e-11 * abs(float(J))
    gq = np.moveaxis(Q0.grad.numpy(), 0, -1)
    gp = np.moveaxis(P0.grad.numpy(), 0, -1)
    for name, a, b in (("q", qb, gq), ("prim", pb, gp)):
        for m in range(a.shape[-1]):
            den = float(np.max(np.abs(b[..., m])))
            e = float(np.max(np.abs(a[..., m] - b[..., m])))
            print(f"  {name}[{m}] max|ref| {den:.3e} err {e:.3e}")
            assert e <= 1e-9 * max(den, float(np.max(np.abs(b)))), (name, m, e, den)
    s.close()


def at_most_one_halo(p):
    """Mask of the haloed nodes outside at most one axis' range 0..n (the
    edges and corners of the halo belong to no state)."""
    ext = [k + 1 for k in (p.im, p.jm, p.km)]
    out = [np.zeros(e + 2 * HM, dtype=int) for e in ext]
    for o, e in zip(out, ext):
        o[:HM] = 1
        o[HM + e:] = 1
    return (out[0][:, None, None] + out[1][None, :, None] + out[2][None, None, :]) <= 1


def test_record_then_reverse():
    """astr_adjoint_record reaches the state of the plain steps bit for bit
    (then qswap); astr_adjoint_reverse gives what astr_adjoint_enstrophy
    gives, with other work on the context in between, and leaves the
    recorded final state; reversing before a record is an error."""
    dims = (16, 18, 20)
    print(f"case record/reverse dims={dims}", flush=True)
    p = ag.default_params(dims, mach=0.3, reynolds=173.0)
    st = ag.hit_state(p, 3)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))

    def up():
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])

    with pytest.raises(ag.AstrError):
        s.adjoint_reverse()
    nsteps, dt, ckpt = 5, 0.004, 2
    up()
    s.steploop(nsteps, dt)
    s.qswap()
    ref = s.download_state()
    up()
    J = s.adjoint_record(nsteps, dt, ckpt)
    got = s.download_state()
    mask = at_most_one_halo(p)
    for k in ("q", "rho", "vel", "prs", "tmp"):
        a, b = ref[k], got[k]
        assert np.array_equal(a[mask], b[mask]), k
    up()
    s.steploop(1, dt)
    qb, pb, ms = s.adjoint_reverse()
    after = s.download_state()
    for k in ("q", "rho", "vel", "prs", "tmp"):
        assert np.array_equal(ref[k][mask], after[k][mask]), k
    up()
    J2, qb2, pb2, _ = s.adjoint_enstrophy(nsteps, dt, ckpt)
    print(f"  J {J:.15e} {J2:.15e}; reversal {ms:.2f} ms")
    assert J == J2
    assert np.array_equal(qb, qb2) and np.array_equal(pb, pb2)
    s.close()


def test_adjoint_against_central_differences():
    """<dJ/du0, v> against (J(u0 + eps v) - J(u0 - eps v)) / (2 eps) on the
    GPU, eps = 1e-6, along random Gaussian directions."""
    import torch

    dims = (16, 16, 16)
    p, box, u0, to_host = adjoint_case(dims, seed=5)
    dt, nsteps = 0.004, 3
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))

    def run(u, grad):
        Q0, P0 = box.initial(u)
        h = to_host(Q0, P0)
        s.upload_state(h["q"], h["rho"], h["vel"], h["prs"], h["tmp"])
        J, qb, pb, _ = s.adjoint_enstrophy(nsteps, dt, 2)
        if not grad:
            return J
        # chain through the initial state's construction from u0
        x = u.clone().requires_grad_(True)
        Q, P = box.initial(x)
        Q.backward(torch.tensor(np.moveaxis(qb, -1, 0).copy()), retain_graph=True)
        P.backward(torch.tensor(np.moveaxis(pb, -1, 0).copy()))
        return J, x.grad.numpy()

    J0, g = run(u0, True)
    rng = np.random.default_rng(11)
    eps = 1e-6
    for k in range(3):
        v = rng.standard_normal(tuple(u0.shape))
        jp = run(u0 + eps * torch.tensor(v), False)
        jm = run(u0 - eps * torch.tensor(v), False)
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(g * v))
        print(f"  direction {k}: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)
    s.close()


def test_autograd_function_gives_the_velocity_gradient():
    """EnstrophyAfterSteps (forward: the plain GPU steps, backward: the GPU
    adjoint) composed with velocity_state: dJ/du0 against PyTorch-CPU
    autograd of the restated forward."""
    import torch

    from astr_gpu_amd.autograd import EnstrophyAfterSteps, velocity_state

    dims = (16, 18, 20)
    p, box, u0, _ = adjoint_case(dims, seed=9)
    dt, nsteps = 0.004, 2
    x = u0.clone().requires_grad_(True)
    Jref = box.objective(x, nsteps, dt)
    Jref.backward()
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    y = u0.clone().requires_grad_(True)
    q0, pr0 = velocity_state(p, y)
    J = EnstrophyAfterSteps.apply(q0, pr0, s, nsteps, dt, 1)
    J.backward()
    print(f"  J {float(J):.15e} ref {float(Jref.detach()):.15e}")
    assert abs(float(J) - float(Jref.detach())) <= 1e-11 * abs(float(Jref.detach()))
    e = float((y.grad - x.grad).abs().max() / x.grad.abs().max())
    print(f"  dJ/du0 rel err {e:.2e}")
    assert e <= 1e-9
    s.close()
