"""Throughput of the HIP solver on the Taylor-Green box (FP64, RK3).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]

Workload: 256^3 periodic Taylor-Green box (Re 1600, Ma 0.1, 643c compact
derivatives, alpha 0.49 compact filter every stage, RK3, dt 0.004), the
whole domain resident in one GPU's HBM.  A step is one full RK3 step (three
stages of filter, qswap, right-hand side, update).  The metric counts unique
periodic grid points (n^3) advanced one RK3 step per second.

Other workloads (--workload): riemann2d, RK3 steps of the 2-D four-quadrant
Riemann problem at 2048^2 with the explicit upwind scheme and the example's
boundary conditions; ensemble, 48 periodic turbulence boxes of mixed sizes
advanced together against looping the plain solver over them; adjoint, the
gradient of the enstrophy after 40 RK3 steps of a 128^3 turbulence box by
the discrete adjoint, its cost over the forward's and its agreement with
central differences.

Prints one JSON line.  roofline: the dominant kernel family, its algorithmic
HBM bytes per launch over its average launch time (HIP events on the
solver's stream, over the timed region).  cpu_baseline: the checker's C
restatement (oracle/, OpenMP over lines) on the host cores, one RK3 step at
the same size.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic doubles per interior node per launch of each kernel family:
# the fields a launch must read and write once (averaged over its x, y, z
# launches where they differ).
FAMILY_DOUBLES = {
    "filter": 10.0,   # q(1:5) in and out, per axis
    "qswap": None,    # boundary planes only
    "grad": 8.0,      # u, v, w, T in; dvel(.,d), dtmp(d) out
    # dvel 9, dtmp 3, T, u, v, w, q 5, p, J in; the axis fluxes G (3 x 5) out
    "gflux": 38.0,
    "stats": None,    # once per step (14 doubles per node), not per stage
    # G(d,1:5) in; x: qrhs out (10); y: qrhs in and out (15); z: qrhs, J,
    # q, qsave in, q out (26; stage 1: qsave out instead of in)
    "flux": (10.0 + 15.0 + 26.0) / 3.0,
    "prim": 11.0,     # q 5 in; rho, u, v, w, p, T out
}
FAMILY_KERNEL = {"filter": "k_sweep<L,0> (filter)", "grad": "k_sweep<L,1> (gradient)",
                 "gflux": "k_gflux", "flux": "k_sweep<L,2> (flux derivatives, RK3 in z)",
                 "prim": "k_prim"}
FAMILY_LAUNCHES = {"filter": 3, "qswap": 2, "grad": 3, "gflux": 1, "stats": 0, "flux": 3,
                   "prim": 1}
FAMILY_NAMES = ["filter", "qswap", "grad", "gflux", "stats", "flux", "prim"]
STATS_DOUBLES = 14.0
# whole-step algorithmic bytes per node: three stages of the kernels above
# plus the statistics once
ALGO_BYTES_PER_PT_STEP = 8.0 * (3.0 * sum(FAMILY_DOUBLES[f] * FAMILY_LAUNCHES[f]
                                          for f in FAMILY_NAMES if FAMILY_DOUBLES[f])
                                + STATS_DOUBLES)


def measured_traffic(family: str, n: int):
    """HBM bytes per launch of a kernel family from the committed rocprof
    PMC summary (profiles/tgv256_traffic.json), for the 256^3 workload."""
    path = os.path.join(ROOT, "profiles", "tgv256_traffic.json")
    if n != 256 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_launch"].get(family)


CPU_SAMPLE_S = 10.0  # the least CPU time a baseline sample runs

# box size, timed steps and warm-up steps of each workload
WORKLOAD_DEFAULTS = {"tgv": (256, 20, 3), "riemann2d": (2048, 20, 3), "adjoint": (128, 20, 3),
                     "ensemble": (None, 10, 1)}


def cpu_baseline(n: int, dt: float):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    b = pyoracle.OracleBox(n, variant=0)
    b.init_tgv()
    # whole RK3 steps until at least CPU_SAMPLE_S have passed
    steps = 0
    t0 = time.perf_counter()
    while steps == 0 or time.perf_counter() - t0 < CPU_SAMPLE_S:
        b.step(dt)
        steps += 1
    el = time.perf_counter() - t0
    thr = pyoracle.oracle_lib().orc_omp_threads()
    b.close()
    return {"value": steps * n ** 3 / el, "unit": "grid-point-RK3-steps/s", "cores": int(thr),
            "kind": "port",
            "sample": f"{steps} RK3 steps of the {n}^3 Taylor-Green box by the C restatement "
                      f"(OpenMP, {thr} threads), {el:.1f} s"}


FP64_PEAK_TFLOPS = 78.6


def measured_flops(n: int):
    """FP64 flops and HBM bytes per upwind interface-flux launch from the
    committed rocprof PMC summary (profiles/riemann2d_upwind.json), for the
    2048^2 workload."""
    path = os.path.join(ROOT, "profiles", "riemann2d_upwind.json")
    if n != 2048 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d.get("fp64_flops_per_launch"), d.get("traffic_bytes_per_launch")


RIEMANN_FACES = (0, 0, 0, 0, 1, 1)  # the example's input.riemann2d bctype


def riemann_params(n: int):
    import astr_gpu_amd as ag

    # one domain bounded at both ends on i and j (ntype 4), k absent
    return ag.default_params((n, n, 0), npdci=4, npdcj=4, npdck=3, convection=1,
                             deriv_explicit=1, recon_schem=3, lchardecomp=1, lfilter=0,
                             diffterm=0, mach=1.0)


def cpu_baseline_riemann(n_small: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    import astr_gpu_amd as ag

    p = riemann_params(n_small)
    st = ag.riemann2d_state(p)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    dt = ag.cfl_dt(p, st, dxi, 0.5)
    box = pyoracle.UwdBox((n_small, n_small, 0), npdc=(4, 4, 3), recon_schem=3,
                          lchardecomp=True, mach=1.0)
    dx9 = dxi.reshape(dxi.shape[:3] + (9,), order="F")
    # one step, then as many more as bring the sample to CPU_SAMPLE_S
    t0 = time.perf_counter()
    o = box.rk3_steps(st, dx9, jac, dt, nsteps=1, bctype=RIEMANN_FACES, udf=1)
    el1 = time.perf_counter() - t0
    more = max(0, math.ceil((CPU_SAMPLE_S - el1) / el1))
    if more:
        box.rk3_steps(o, dx9, jac, dt, nsteps=more, time=dt, bctype=RIEMANN_FACES, udf=1)
    el = time.perf_counter() - t0
    steps = 1 + more
    return {"value": steps * (n_small + 1) ** 2 / el, "unit": "grid-point-RK3-steps/s",
            "cores": 1, "kind": "port",
            "sample": f"{steps} RK3 steps of the {n_small}^2 Riemann box by the C restatement "
                      f"(one thread), {el:.1f} s"}


def bench_riemann(args):
    """RK3 steps of the 2-D four-quadrant Riemann problem on an n x n box of
    [0,1]^2 (explicit upwind MP7 with characteristic projection where the
    Ducros sensor flags a shock, explicit 6th-order gradient, the example's
    boundary condition on the four faces, dt from CFL 0.5): grid points
    ((n+1)^2, every node is unique on a bounded box) advanced per second.
    --t-end also runs the problem from t = 0 to that time and reports the
    wall time."""
    import astr_gpu_amd as ag

    n = args.n
    p = riemann_params(n)
    s = ag.AstrSolver(p, device=0)
    jac, dxi = ag.box_metrics(p, lengths=(1.0, 1.0, 1.0))
    s.upload_geom(jac, dxi)
    st = ag.riemann2d_state(p)
    dt = ag.cfl_dt(p, st, dxi, 0.5)
    s.set_bc(RIEMANN_FACES, ag.UDF_RIEMANN2D)

    def restart():
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        s.time = 0.0

    restart()
    s.steploop(args.warmup, dt)
    s.synchronize()
    s.set_timing(True)
    s.steploop(args.steps, dt)
    fam = s.timing()
    s.set_timing(False)
    restart()
    s.synchronize()
    t0 = time.perf_counter()
    s.steploop(args.steps, dt)
    s.synchronize()
    el = time.perf_counter() - t0
    pts = float(n + 1) ** 2
    value = pts * args.steps / el
    avg_ms = fam[5] / (6 * args.steps)  # two interface-flux sweeps per stage
    flops, traffic = measured_flops(n)
    out = {
        "metric": "FP64 grid-point-RK3-steps/s, 2-D Riemann (explicit upwind)",
        "value": value, "unit": "grid-point-RK3-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (four-quadrant Riemann initial field of riem2dini)",
        "config": {"workload": f"2-D four-quadrant Riemann {n}^2 on [0,1]^2, recon MP7, "
                               f"characteristic projection, Ducros sensor, explicit 643e "
                               f"gradient, udf_bc on four faces, dt {dt:.4e} (CFL 0.5)",
                   "n": n, "parallelism": "single GPU"},
        "family_ms_per_step": {"boucon+qswap": fam[1] / args.steps, "grad": fam[2] / args.steps,
                               "sensor": fam[3] / args.steps,
                               "interface_flux": fam[5] / args.steps,
                               "rk_update": fam[6] / args.steps, "total": fam[7] / args.steps},
        "roofline": None,
    }
    if flops:
        achieved = flops / (avg_ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                           "traffic": traffic,
                           "kernel": "k_uwd (interface fluxes), 2 launches/stage",
                           "flops_per_launch": flops, "avg_launch_ms": avg_ms}
    if args.t_end:
        restart()
        nsteps = 0
        t = 0.0
        while t < args.t_end - 1e-12:
            t = t + dt
            nsteps += 1
        s.synchronize()
        t0 = time.perf_counter()
        s.steploop(nsteps, dt)
        s.synchronize()
        run = time.perf_counter() - t0
        fin = s.download_state()
        rho = ag.interior(fin["rho"], p)
        prs = ag.interior(fin["prs"], p)
        out["run"] = {"t_end": s.time, "steps": nsteps, "wall_s": run,
                      "rho_min": float(rho.min()), "rho_max": float(rho.max()),
                      "p_min": float(prs.min()), "p_max": float(prs.max()),
                      "finite": bool(np.isfinite(fin["q"]).all())}
    s.close()
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline_riemann(1024)
    print(json.dumps(out))


def bench_adjoint(args):
    """Discrete adjoint on the decaying-turbulence box (n^3, HIT spectrum,
    seeded; Mt 0.3, Re 173): J = enstrophy after nsteps RK3 steps, dJ/du0
    by the GPU adjoint (checkpoints every ckpt steps), checked against
    central differences (eps 1e-6) along 8 seeded Gaussian directions.  The
    metric is the device time of the reversal (the backward of the autograd
    binding) over the time of the plain forward steps (lower is better); the
    whole gradient with its recorded forward is reported beside it."""
    import torch

    import astr_gpu_amd as ag
    from astr_gpu_amd.autograd import velocity_state

    n, nsteps, ckpt, dt = args.n, args.nsteps, args.ckpt, 0.002
    p = ag.default_params(n, mach=0.3, reynolds=173.0)
    s = ag.AstrSolver(p, device=0)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.hit_state(p, args.seed)
    u0 = np.moveaxis(ag.interior(st["vel"], p)[:-1, :-1, :-1], -1, 0).copy()

    def upload(state):
        s.upload_state(state["q"], state["rho"], state["vel"], state["prs"], state["tmp"])

    def objective(u):
        upload(ag.state_from_velocity(p, u))
        s.steploop(nsteps, dt)
        s.qswap()
        s.rhscal()
        return s.statcal()[1]

    # warm-up of every kernel involved
    upload(st)
    s.adjoint_enstrophy(min(nsteps, 2), dt, ckpt)
    # plain forward
    reps = max(1, args.steps // 10)
    fwd = []
    for _ in range(reps):
        upload(st)
        s.synchronize()
        t0 = time.perf_counter()
        s.steploop(nsteps, dt)
        s.synchronize()
        fwd.append(time.perf_counter() - t0)
    t_fwd = min(fwd)

    # gradient: the recorded forward (astr_adjoint_record: checkpoints every
    # k steps, the last segment's stages stored), then the reversal
    # (astr_adjoint_reverse: the earlier segments recomputed, every stage
    # reversed; device time by HIP events, and the wall time with the
    # download of the two haloed gradients to host memory)
    def gradient(k):
        best, res = None, None
        for _ in range(reps):
            upload(st)
            s.synchronize()
            t0 = time.perf_counter()
            J = s.adjoint_record(nsteps, dt, k)
            s.synchronize()
            t1 = time.perf_counter()
            qb, pb, ms = s.adjoint_reverse()
            t2 = time.perf_counter()
            cur = (t1 - t0, 1e-3 * ms, t2 - t1)
            if best is None or cur[1] < best[1]:
                best, res = cur, (J, qb, pb)
        return best, res

    (rec_all, rev_all, _), _ = gradient(nsteps)
    (t_rec, t_rev, t_rev_wall), (J, qb, pb) = gradient(ckpt)
    # dJ/du0 through the state construction
    x = torch.tensor(u0, requires_grad=True)
    q0, pr0 = velocity_state(p, x)
    torch.autograd.backward([q0, pr0], [torch.from_numpy(qb), torch.from_numpy(pb)])
    g = x.grad.numpy()
    rng = np.random.default_rng(args.seed)
    errs = {1e-6: [], 1e-4: []}
    for _ in range(args.fd_dirs):
        v = rng.standard_normal(u0.shape)
        ad = float(np.sum(g * v))
        for eps in errs:
            fd = (objective(u0 + eps * v) - objective(u0 - eps * v)) / (2 * eps)
            errs[eps].append(abs(fd - ad) / abs(ad))
    s.close()
    out = {
        "metric": "adjoint/forward cost ratio (reversal of the recorded steps over the plain "
                  "forward steps)",
        "value": t_rev / t_fwd, "unit": "ratio", "n_gpus": 1, "steps": nsteps,
        "warmup": 1, "ms_per_step": 1e3 * t_fwd / nsteps, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded decaying isotropic turbulence)",
        "config": {"workload": f"discrete adjoint, HIT {n}^3 Mt 0.3 Re 173 seed {args.seed}, "
                               f"{nsteps} RK3 steps dt {dt}, checkpoints every {ckpt} steps, "
                               f"J = enstrophy of the final state",
                   "n": n, "parallelism": "single GPU"},
        "forward_s": t_fwd, "recorded_forward_s": t_rec, "reversal_s": t_rev,
        "reversal_with_download_s": t_rev_wall,
        "gradient_total_ratio": (t_rec + t_rev) / t_fwd,
        "ratio_all_stages_kept": rev_all / t_fwd,
        "gradient_total_ratio_all_stages_kept": (rec_all + rev_all) / t_fwd,
        "J": J, "fd_max_rel_err": max(errs[1e-6], default=None), "fd_eps": 1e-6,
        "fd_directions": args.fd_dirs, "fd_max_rel_err_eps_1e-4": max(errs[1e-4], default=None),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


ENSEMBLE_SIZES = (64, 96, 128, 160, 192)


def ensemble_dims(nbox: int, seed: int):
    rng = __import__("numpy").random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(ENSEMBLE_SIZES, 3)) for _ in range(nbox)]


def bench_ensemble(args):
    """48 independent periodic boxes of decaying isotropic turbulence (sizes
    drawn from {64,...,192} per axis, seeded), the central scheme with
    filter and viscous terms: aggregate grid-point-RK3-steps/s of the
    batched graph against looping the plain solver over the boxes."""
    import astr_gpu_amd as ag

    dims = ensemble_dims(args.nbox, args.seed)
    dt = 0.002
    # Re_lambda ~ 50 for the k^4 exp(-2 (k/4)^2) spectrum at sqrt(<u.u>) = 1:
    # lambda = 0.5, u' = 3^-1/2, so Re = 50 / (0.5 * 3^-1/2) = 173
    solvers = []
    for i, d in enumerate(dims):
        p = ag.default_params(d, mach=0.3, reynolds=173.0)
        s = ag.AstrSolver(p, device=0)
        s.upload_geom(*ag.box_metrics(p))
        st = ag.hit_state(p, args.seed * 1000 + i)
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        solvers.append(s)
    pts = float(sum(a * b * c for a, b, c in dims))
    # plain solver, box after box
    for s in solvers:
        s.steploop(args.warmup, dt)
    t0 = time.perf_counter()
    for s in solvers:
        s.steploop(args.steps, dt)
    el_loop = time.perf_counter() - t0
    # batched: one graph per step over all boxes, and the same launches made
    # directly on the members' streams
    b = ag.AstrBatch(solvers)
    b.rk3_steps(args.warmup, dt)
    t0 = time.perf_counter()
    b.rk3_steps(args.steps, dt)
    el_graph = time.perf_counter() - t0
    b.set_mode(1)
    b.rk3_steps(args.warmup, dt)
    t0 = time.perf_counter()
    b.rk3_steps(args.steps, dt)
    el_streams = time.perf_counter() - t0
    b.close()
    el = min(el_graph, el_streams)
    for s in solvers:
        s.close()
    value = pts * args.steps / el
    loop_value = pts * args.steps / el_loop
    out = {
        "metric": "FP64 grid-point-RK3-steps/s, batched ensemble",
        "value": value, "unit": "grid-point-RK3-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded decaying isotropic turbulence)",
        "config": {"workload": f"{args.nbox} periodic boxes, sizes per axis from "
                               f"{list(ENSEMBLE_SIZES)} (seed {args.seed}), HIT Mt 0.3, "
                               f"Re 173, 643c, filter 0.49, RK3 dt {dt}",
                   "boxes": [list(d) for d in dims], "points": pts,
                   "parallelism": "single GPU"},
        "loop_value": loop_value,
        "batched_over_loop": value / loop_value,
        "graph_value": pts * args.steps / el_graph,
        "streams_value": pts * args.steps / el_streams,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 20 (ensemble: 10)")
    ap.add_argument("--warmup", type=int, default=None, help="default 3 (ensemble: 1)")
    ap.add_argument("--n", type=int, default=None,
                    help="box size; default 256 (tgv), 2048 (riemann2d), 128 (adjoint)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["tgv", "riemann2d", "ensemble", "adjoint"],
                    default="tgv",
                    help="tgv: the headline RK3 step; riemann2d: RK3 steps of the 2-D "
                         "Riemann problem (upwind, boundary conditions); ensemble: batched "
                         "boxes against looping the plain solver")
    ap.add_argument("--nsteps", type=int, default=40, help="adjoint: RK3 steps of the objective")
    ap.add_argument("--ckpt", type=int, default=8, help="adjoint: steps between kept states")
    ap.add_argument("--fd-dirs", type=int, default=8,
                    help="adjoint: directions of the central-difference check (0: none)")
    ap.add_argument("--t-end", type=float, default=0.0,
                    help="riemann2d: also run the problem from t = 0 to this time (0.3 in "
                         "the benchmark) and report its wall time")
    ap.add_argument("--nbox", type=int, default=48)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the solver keeps the whole box on one device "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    # each workload's own sizes where the command line leaves them open
    n0, steps0, warmup0 = WORKLOAD_DEFAULTS[args.workload]
    args.n = n0 if args.n is None else args.n
    args.steps = steps0 if args.steps is None else args.steps
    args.warmup = warmup0 if args.warmup is None else args.warmup
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "riemann2d":
        bench_riemann(args)
        return

    import astr_gpu_amd as ag

    n, dt = args.n, 0.004
    p = ag.default_params(n)
    s = ag.AstrSolver(p, device=0)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.tgv_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    del st

    s.steploop(args.warmup, dt)
    # timed region: K steps, kernel families timed with events on the stream
    s.set_timing(True)
    t0 = time.perf_counter()
    s.steploop(args.steps, dt)   # synchronizes before returning
    el = time.perf_counter() - t0
    fam = s.timing()
    s.set_timing(False)
    # untimed-by-events pass for the wall clock without event overhead
    t0 = time.perf_counter()
    s.steploop(args.steps, dt)
    el_noev = time.perf_counter() - t0
    el = min(el, el_noev)
    copy_ms, copy_bytes = s.bench_copy(20)
    s.close()
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9

    pts = float(n) ** 3
    value = pts * args.steps / el
    ms_per_step = 1e3 * el / args.steps
    nodes = float(n + 1) ** 3
    stages = 3 * args.steps
    fam_ms = {nm: float(fam[i]) for i, nm in enumerate(FAMILY_NAMES)}
    fam_ms["total"] = float(fam[7])
    dom = max((k for k in FAMILY_NAMES if FAMILY_DOUBLES[k]), key=lambda k: fam_ms[k])
    launches = FAMILY_LAUNCHES[dom] * stages
    avg_s = fam_ms[dom] / 1e3 / launches
    algo = FAMILY_DOUBLES[dom] * 8.0 * nodes
    achieved = algo / avg_s / 1e9
    out = {
        "metric": "FP64 grid-point-RK3-steps/s at 256^3 (plain)",
        "value": value,
        "unit": "grid-point-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green initial field)",
        "config": {"workload": f"Taylor-Green vortex {n}^3 periodic, Re 1600, Ma 0.1, 643c, "
                               f"filter alpha 0.49, RK3 dt 0.004", "n": n,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, n),
                     "kernel": f"{FAMILY_KERNEL[dom]}, {FAMILY_LAUNCHES[dom]} launches/stage",
                     "algo_bytes_per_launch": algo, "avg_launch_ms": avg_s * 1e3},
        "step_hbm_frac": value * ALGO_BYTES_PER_PT_STEP / (HBM_PEAK_GBS * 1e9),
        "copy_gbs": copy_gbs,
        "step_copy_frac": value * ALGO_BYTES_PER_PT_STEP / (copy_gbs * 1e9),
        "family_ms_per_step": {k: v / args.steps for k, v in fam_ms.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, dt)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
