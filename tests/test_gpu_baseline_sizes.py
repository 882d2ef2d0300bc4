"""The HIP path at the sizes the benchmark workloads run.

Line sweeps are instantiated per rows-per-lane L = ceil((n + 7) / 32): the
256^3 Taylor-Green box runs every sweep at L = 9, the 128^3 adjoint box the
transposed sweeps at L = 5, and the 48-box ensemble L = 3..7 (axes of 64 to
192 cells).  These tests drive exactly those instantiations:

  * 256^3 Taylor-Green, 100 RK3 steps (dt 0.001, the acoustic CFL number of
    the 64^3 box at dt 0.004) against the checker's series written by
    tests/golden/make_oracle_series.py (KE and enstrophy
    within 1e-10 relative at every step, primitive fields within 1e-9
    absolute at the probes and along three whole lines), and one stage's
    pieces plus one whole step at 256^3 against the checker, field by field;
  * the 128^3 adjoint of the enstrophy after 40 RK3 steps, checkpoints every
    8 steps (five segments, four recomputed), against central differences
    (eps 1e-6) along 8 seeded Gaussian directions, bar 1e-6 relative; and
    the multi-segment recompute against PyTorch-CPU autograd of the
    restated forward;
  * the seed-1 ensemble of 48 boxes, 50 batched RK3 steps (25 as one graph
    over the members' streams, 25 through the autotuned line-queue engine),
    every member bit for bit equal to the same steps taken box by box, and
    the smallest and the largest member against the checker for 2 steps.
"""
import os

import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle
from conftest import GOLDEN
from test_gpu_parity import PRIMS, compare_state

pytestmark = pytest.mark.gpu
HM = 5


def say(msg):
    print(msg, flush=True)


def upload(s, st):
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])


# ------------------------------------------------------- Taylor-Green 256^3

def test_tgv256_100_steps_against_checker_series():
    """The headline workload: k_sweep<9, PASS> of every pass, x and y/z
    forms, and the z sweep's RK3 combine, over 100 steps."""
    path = os.path.join(GOLDEN, "tgv256_main.npz")
    G = np.load(path)
    n, dt, steps = int(G["n"]), float(G["dt"]), int(G["steps"])
    assert n == 256 and steps == 100
    say("case tgv256 100 steps: checker initial field")
    ob = pyoracle.OracleBox(n, variant=0)
    ob.init_tgv()
    st = ob.state_haloed()
    ob.close()
    p = ag.default_params(n)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    upload(s, st)
    del st
    say("case tgv256 100 steps: stepping")
    stats = np.asarray(s.steploop(steps, dt, stats=True))
    ref = G["stats"]
    for c, nm in enumerate(("KE", "enstrophy", "dissipation")):
        e = float(np.max(np.abs(stats[:, c] - ref[:, c]) / np.abs(ref[:, c])))
        say(f"  {nm} max rel err over {steps} steps {e:.2e}")
        assert e < 1e-10, (nm, e)
    out = s.download_state()
    s.close()
    st_ = n // 8
    i0, j0, k0 = (3 * st_ + 1, 5 * st_ + 2, 7 * st_ + 3)
    fields = [("q1", "q", 0), ("q2", "q", 1), ("q3", "q", 2), ("q4", "q", 3), ("q5", "q", 4)]
    fields += [(nm, "vel", c) if c is not None else (nm, nm, None) for nm, c in PRIMS]
    for nm, key, c in fields:
        g = out[key][..., c] if c is not None else out[key]
        g = ag.interior(g, p)
        worst = max(float(np.max(np.abs(g[::st_, ::st_, ::st_] - G["probe_" + nm]))),
                    float(np.max(np.abs(g[:, j0, k0] - G["xline_" + nm]))),
                    float(np.max(np.abs(g[i0, :, k0] - G["yline_" + nm]))),
                    float(np.max(np.abs(g[i0, j0, :] - G["zline_" + nm]))))
        l2 = float(np.sqrt(np.mean(g * g)))
        say(f"  {nm}: probes and lines max abs err {worst:.2e}; rms rel err "
            f"{abs(l2 - float(G['l2_' + nm])) / max(float(G['l2_' + nm]), 1e-300):.2e}")
        assert worst < 1e-9, (nm, worst)


def test_tgv256_stage_pieces_and_one_step_against_checker():
    """filterq, qswap and rhscal at 256^3 (the filter, gradient and flux
    sweeps at L = 9 without the RK3 epilogue), then one whole step from the
    initial field, every node and every halo face qswap writes."""
    n, dt = 256, 0.001
    say("case tgv256 stage pieces")
    ob = pyoracle.OracleBox(n, variant=0)
    ob.init_tgv()
    st = ob.state_haloed()
    p = ag.default_params(n)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    upload(s, st)
    ob.filterq()
    s.filterq()
    say(f"  filterq {compare_state(ob, s, p, 1e-12):.2e}")
    ob.qswap()
    s.qswap()
    say(f"  qswap {compare_state(ob, s, p, 1e-12, halo_faces=True):.2e}")
    ob.rhscal()
    s.rhscal()
    r = s.download_rhs()
    flux_scale = float(np.max(np.abs(ob.get("prs")))) * n / (2 * np.pi)
    for m in range(5):
        e = float(np.max(np.abs(r["qrhs"][..., m] - ob.deriv(12 + m)))) / flux_scale
        say(f"  qrhs{m + 1}: {e:.2e} of the flux scale")
        assert e < 1e-13
    xi = n / (2 * np.pi)
    fs = max(float(np.max(np.abs(ob.get(nm)))) for nm in "uvw") * xi
    for m in range(3):
        for d in range(3):
            e = float(np.max(np.abs(r["dvel"][..., m, d] - ob.deriv(3 * m + d)))) / fs
            assert e < 1e-13, ("dvel", m, d, e)
    fs = float(np.max(np.abs(ob.get("tmp")))) * xi
    for d in range(3):
        e = float(np.max(np.abs(r["dtmp"][..., d] - ob.deriv(9 + d)))) / fs
        assert e < 1e-13, ("dtmp", d, e)
    del r
    ob.close()
    say("case tgv256 one whole step")
    ob = pyoracle.OracleBox(n, variant=0)
    ob.init_tgv()
    upload(s, st)
    del st
    so = ob.step(dt, stats=True)
    sg = s.time_integration_rk(dt, stats=True)
    # the checker adds the 256^3 terms of each statistic one by one in the
    # reference's loop order (src/statistic.F90:870-1042), which at this
    # size rounds at ~1e-12 relative; the GPU sums in double-double.  The
    # statistics bar is therefore the whole-run bar, 1e-10.
    for a, b in zip(sg, so):
        assert abs(a - b) <= 1e-10 * abs(b), (sg, so)
    say(f"  one step {compare_state(ob, s, p, 1e-12, halo_faces=True, vector_scale=True):.2e}")
    s.close()
    ob.close()


# ------------------------------------------------------------ adjoint 128^3

def test_adjoint128_40_steps_ckpt8_against_central_differences():
    """The adjoint workload at its size: dJ/du0 of the enstrophy after 40
    RK3 steps (dt 0.002) of the seed-1 128^3 turbulence box, checkpoints
    every 8 steps, against (J(u0 + eps v) - J(u0 - eps v)) / (2 eps) of
    the GPU forward, eps = 1e-6, along 8 Gaussian directions of a
    generator seeded with 1; bar 1e-6 relative."""
    import torch

    from astr_gpu_amd.autograd import velocity_state

    n, nsteps, ckpt, dt, seed = 128, 40, 8, 0.002, 1
    say("case adjoint128 40 steps ckpt 8")
    p = ag.default_params(n, mach=0.3, reynolds=173.0)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.hit_state(p, seed)
    u0 = np.moveaxis(ag.interior(st["vel"], p)[:-1, :-1, :-1], -1, 0).copy()
    upload(s, st)
    J, qb, pb, ms = s.adjoint_enstrophy(nsteps, dt, ckpt)
    say(f"  J {J:.15e}; recorded forward {ms[0]:.1f} ms, reversal {ms[1]:.1f} ms")
    x = torch.tensor(u0, requires_grad=True)
    q0, pr0 = velocity_state(p, x)
    torch.autograd.backward([q0, pr0], [torch.from_numpy(qb), torch.from_numpy(pb)])
    g = x.grad.numpy()

    def objective(u):
        upload(s, ag.state_from_velocity(p, u))
        s.steploop(nsteps, dt)
        s.qswap()
        s.rhscal()
        return s.statcal()[1]

    j0 = objective(u0)
    say(f"  J of the plain forward {j0:.15e}")
    assert abs(j0 - J) <= 1e-12 * abs(J)
    rng = np.random.default_rng(seed)
    eps = 1e-6
    worst = 0.0
    for k in range(8):
        v = rng.standard_normal(u0.shape)
        ad = float(np.sum(g * v))
        fd = (objective(u0 + eps * v) - objective(u0 - eps * v)) / (2 * eps)
        e = abs(fd - ad) / abs(ad)
        worst = max(worst, e)
        say(f"  direction {k}: fd {fd:.12e} adjoint {ad:.12e} rel {e:.2e}")
    s.close()
    assert worst <= 1e-6, worst


@pytest.mark.parametrize("nsteps,ckpt", [(12, 4), (10, 4)])
def test_adjoint_segments_recomputed_against_torch_autograd(nsteps, ckpt):
    """Three segments of 4 steps (two recomputed from their kept states),
    and 4 + 4 + 2 (a last segment shorter than ckpt), on a 24^3 box:
    dJ/d(state) against PyTorch-CPU autograd of the restated forward."""
    import torch
    from torch_forward import Box, to_host

    dims, dt = (24, 24, 24), 0.004
    say(f"case adjoint segments dims={dims} nsteps={nsteps} ckpt={ckpt}")
    p = ag.default_params(dims, mach=0.3, reynolds=173.0)
    st = ag.hit_state(p, 3)
    u0 = torch.tensor(np.moveaxis(ag.interior(st["vel"], p)[:-1, :-1, :-1], -1, 0).copy())
    box = Box(dims, reynolds=p.reynolds, mach=p.mach)
    Q0, P0 = box.initial(u0)
    Q0 = Q0.detach().requires_grad_(True)
    P0 = P0.detach().requires_grad_(True)
    J = box.objective_state(Q0, P0, nsteps, dt)
    J.backward()
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    h = to_host(Q0, P0)
    upload(s, h)
    Jg, qb, pb, _ = s.adjoint_enstrophy(nsteps, dt, ckpt)
    say(f"  J gpu {Jg:.15e} torch {float(J.detach()):.15e}")
    assert abs(Jg - float(J)) <= 1e-11 * abs(float(J))
    gq = np.moveaxis(Q0.grad.numpy(), 0, -1)
    gp = np.moveaxis(P0.grad.numpy(), 0, -1)
    for name, a, b in (("q", qb, gq), ("prim", pb, gp)):
        top = float(np.max(np.abs(b)))
        for m in range(a.shape[-1]):
            den = float(np.max(np.abs(b[..., m])))
            e = float(np.max(np.abs(a[..., m] - b[..., m])))
            say(f"  {name}[{m}] max|ref| {den:.3e} err {e:.3e}")
            assert e <= 1e-9 * max(den, top), (name, m, e, den)
    s.close()


# ------------------------------------------------------- ensemble of 48 boxes

def test_ensemble48_batched_equals_box_by_box_and_checker():
    """The seed-1 ensemble (48 boxes, axes from {64, 96, 128, 160, 192}):
    50 batched RK3 steps (one graph per step) leave every member bit for bit
    where 50 plain steps of the same box take it; the smallest and the
    largest member also step twice against the checker (1e-10)."""
    nbox, seed, dt, steps = 48, 1, 0.002, 50
    dims = ag.ensemble_dims(nbox, seed)
    Ls = sorted({(d + 7 + 31) // 32 for dd in dims for d in dd})
    say(f"case ensemble48: rows per lane {Ls}")
    assert Ls == [3, 4, 5, 6, 7]
    params, states, members = [], [], []
    for i, d in enumerate(dims):
        p = ag.default_params(d, mach=0.3, reynolds=173.0)
        st = ag.hit_state(p, ag.ensemble_member_seed(seed, i))
        s = ag.AstrSolver(p)
        s.upload_geom(*ag.box_metrics(p))
        upload(s, st)
        params.append(p)
        states.append(st)
        members.append(s)
        if i % 8 == 7:
            say(f"  {i + 1} members created")
    b = ag.AstrBatch(members)
    b.rk3_steps(steps // 2, dt)
    say(f"  {steps // 2} batched steps (graph over the members' streams) done")
    b.set_mode(2)
    rec = b.autotune(dt)
    say(f"  line-queue tuning: {[(e['kind'], e['L'], e['config']) for e in rec]}")
    b.rk3_steps(steps - steps // 2, dt)
    b.close()
    say(f"  {steps - steps // 2} batched steps (line-queue engine) done")
    for i, (d, p, st, s) in enumerate(zip(dims, params, states, members)):
        got = s.download_state()
        s.close()
        ref = ag.AstrSolver(p)
        ref.upload_geom(*ag.box_metrics(p))
        upload(ref, st)
        ref.steploop(steps, dt)
        want = ref.download_state()
        ref.close()
        for nm in want:
            assert np.array_equal(got[nm], want[nm]), (i, d, nm)
        say(f"  member {i} {d}: bit-equal to box-by-box steps")
    pts = [a * b_ * c for a, b_, c in dims]
    for i in (int(np.argmin(pts)), int(np.argmax(pts))):
        d, p, st = dims[i], params[i], states[i]
        say(f"case ensemble48 member {i} {d} against the checker")
        ob = pyoracle.OracleBox(d, variant=0, reynolds=p.reynolds, mach=p.mach)
        for m in range(5):
            ob.set(f"q{m + 1}", ag.interior(st["q"][..., m], p))
        ob.set("rho", ag.interior(st["rho"], p))
        for c, nm in enumerate("uvw"):
            ob.set(nm, ag.interior(st["vel"][..., c], p))
        ob.set("prs", ag.interior(st["prs"], p))
        ob.set("tmp", ag.interior(st["tmp"], p))
        s = ag.AstrSolver(p)
        s.upload_geom(*ag.box_metrics(p))
        upload(s, st)
        for k in range(2):
            so = ob.step(dt, stats=True)
            sg = s.time_integration_rk(dt, stats=True)
            for a, c in zip(sg, so):
                assert abs(a - c) <= 1e-10 * abs(c), (i, k, sg, so)
        say(f"  2 steps {compare_state(ob, s, p, 1e-10, halo_faces=True):.2e}")
        s.close()
        ob.close()


# ------------------------------------------------ 2-D Riemann problem to t = 0.3

def riemann_case(n):
    import bench

    p = bench.riemann_params(n)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    st = ag.riemann2d_state(p)
    dt = ag.cfl_dt(p, st, dxi, 0.5)
    nsteps, t = 0, 0.0
    while t < 0.3 - 1e-12:
        t, nsteps = t + dt, nsteps + 1
    s = ag.AstrSolver(p)
    s.upload_geom(jac, dxi)
    s.set_bc(bench.RIEMANN_FACES, ag.UDF_RIEMANN2D)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.time = 0.0
    return p, s, st, jac, dxi, dt, nsteps, bench.RIEMANN_FACES


def test_riemann2d_256_to_t03_bit_equal_to_the_checker():
    """The whole four-quadrant problem (MP7, characteristic projection,
    Ducros sensor, the example's boundary condition, CFL 0.5) to t = 0.3 on
    a 256^2 grid: the GPU's final state equals the checker's bit for bit
    (both evaluate the upwind path without FMA contraction, in the
    reference's order)."""
    p, s, st, jac, dxi, dt, nsteps, faces = riemann_case(256)
    say(f"case riemann2d 256^2 to t = 0.3: {nsteps} steps of {dt:.6e}")
    s.steploop(nsteps, dt)
    g = s.download_state()
    s.close()
    box = pyoracle.UwdBox((256, 256, 0), npdc=(4, 4, 3), recon_schem=3, lchardecomp=True,
                          bfacmpld=0.7, shkcrt=0.01, gamma=p.gamma, mach=p.mach)
    dx9 = dxi.reshape(dxi.shape[:3] + (9,), order="F")
    o = st
    done, t = 0, 0.0
    while done < nsteps:
        k = min(100, nsteps - done)
        # the start time of the chunk as the steps accumulate it (t = t + dt)
        o = box.rk3_steps(o, dx9, jac, dt, nsteps=k, time=t, bctype=faces, udf=1)
        for _ in range(k):
            t = t + dt
        done += k
        say(f"  checker {done}/{nsteps} steps")
    for key in ("q", "rho", "vel", "prs", "tmp"):
        a, b = g[key][:, :, HM], o[key][:, :, HM]
        say(f"  {key}: {'bit-equal' if np.array_equal(a, b) else np.max(np.abs(a - b))}")
        assert np.array_equal(a, b), key


def test_riemann2d_2048_to_t03():
    """The benchmark size: 2048^2 to t = 0.3 (5148 steps).  The state stays
    finite, density and pressure positive and within the range the
    quadrant states allow.  The mirror difference across the diagonal is
    printed, not bounded: the reference's north-west pressure literal is
    single precision (0.30 = 0.3 + 1.2e-8, kept in the initial field and
    the boundary condition), and the slip lines amplify that asymmetry; a
    checker run from an exactly symmetric field stays symmetric to 2e-8
    after 100 steps at 128^2, against 4e-6 with the literal."""
    p, s, st, jac, dxi, dt, nsteps, faces = riemann_case(2048)
    say(f"case riemann2d 2048^2 to t = 0.3: {nsteps} steps")
    s.steploop(nsteps, dt)
    g = s.download_state()
    t_end = s.time
    s.close()
    assert abs(t_end - 0.3) < dt
    rho = ag.interior(g["rho"], p)[:, :, 0]
    prs = ag.interior(g["prs"], p)[:, :, 0]
    u = ag.interior(g["vel"][..., 0], p)[:, :, 0]
    v = ag.interior(g["vel"][..., 1], p)[:, :, 0]
    say(f"  rho [{rho.min():.4f}, {rho.max():.4f}], p [{prs.min():.4f}, {prs.max():.4f}], "
        f"mirror difference rho {np.max(np.abs(rho - rho.T)):.3e}, "
        f"u/v {np.max(np.abs(u - v.T)):.3e}")
    assert np.all(np.isfinite(g["q"]))
    assert rho.min() > 0.1 and rho.max() < 2.0
    assert prs.min() > 0.01 and prs.max() < 2.0
