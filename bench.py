"""Throughput of the HIP solver (FP64, RK3): the three parts of the metric.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
                    [--legs plain,batched,adjoint]

Plain (the value): the 256^3 periodic Taylor-Green box (Re 1600, Ma 0.1,
643c compact derivatives, alpha 0.49 compact filter every stage, RK3 with
dt 0.001, i.e. the acoustic CFL number of the 64^3 box at dt 0.004), the
whole domain resident in one GPU's HBM, the statistics of every step
computed as rkfirst does.  A step is one full RK3 step (three stages of
filter, qswap, right-hand side, update).  The metric counts unique periodic
grid points (n^3) advanced one RK3 step per second: K steps between two HIP
events on the solver's stream, median of 5 runs.

Batched (the "batched" object): 48 periodic turbulence boxes of mixed sizes
(seed 1) advanced together, against looping the plain solver over them.
Adjoint (the "adjoint" object): the gradient of the enstrophy after 40 RK3
steps of a 128^3 turbulence box by the discrete adjoint, its device-time
cost over the plain forward's and its agreement with central differences.

Other workloads (--workload): riemann2d, RK3 steps of the 2-D four-quadrant
Riemann problem at 2048^2 with the explicit upwind scheme and the example's
boundary conditions; ensemble and adjoint, those legs alone.

Prints one JSON line.  roofline: the dominant kernel family, its algorithmic
HBM bytes per launch over its average launch time (HIP events on the
solver's stream).  step_roofline: the whole step against the reference's
pass structure, 4 224 B per grid-point-RK3-step.  cpu_baseline: the
checker's C restatement (oracle/, OpenMP over lines) on the host cores,
whole RK3 steps of the same box for at least 10 s.
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
    # dvel 9, dtmp 3, T, u, v, w, q 5, p in; the axis fluxes G (3 x 5) out
    # (a grid whose Jacobian varies also reads J; the box's is one scalar)
    "gflux": 37.0,
    "stats": None,    # inside stage 1's k_gflux: one more load (rho) per step
    # G(d,1:5) in; x: qrhs out (10); y: qrhs in and out (15); z: qrhs, q,
    # qsave in, q out (25 at stages 2 and 3; 20 at stage 1, where q and
    # qsave swap roles instead of storing qsave; J a scalar)
    "flux": (10.0 + 15.0 + (20.0 + 25.0 + 25.0) / 3.0) / 3.0,
    # q 5 in; u, v, w, p, T out, and rho at the last stage (no kernel reads
    # the density in between)
    "prim": (10.0 + 10.0 + 11.0) / 3.0,
}
FAMILY_KERNEL = {"filter": "k_sweep<L,0> (filter)", "grad": "k_sweep<L,1> (gradient)",
                 "gflux": "k_gflux", "flux": "k_sweep<L,2> (flux derivatives, RK3 in z)",
                 "prim": "k_prim"}
FAMILY_LAUNCHES = {"filter": 3, "qswap": 2, "grad": 3, "gflux": 1, "stats": 0, "flux": 3,
                   "prim": 1}
FAMILY_NAMES = ["filter", "qswap", "grad", "gflux", "stats", "flux", "prim"]
STATS_DOUBLES = 1.0  # the statistics' rho, loaded by stage 1's k_gflux
# the reference's own pass structure: 176 doubles per point per stage
REF_BYTES_PER_PT_STEP = 4224.0
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


def cpu_model() -> str:
    """The host CPU's model name (what lscpu prints as "Model name")."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
            "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{steps} RK3 steps (dt {dt:g}) of the {n}^3 Taylor-Green box by the C "
                      f"restatement (OpenMP, {thr} threads), {el:.1f} s"}


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
            "cores": 1, "kind": "port", "cpu_model": cpu_model(),
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
        out["roofline"] = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
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


def adjoint_leg(n=128, nsteps=40, ckpt=8, seed=1, fd_dirs=8, fd_eps=1e-6):
    """Discrete adjoint on the decaying-turbulence box (n^3, HIT spectrum,
    seeded; Mt 0.3, Re 173): J = enstrophy after nsteps RK3 steps (dt
    0.002), dJ/du0 by the GPU adjoint (checkpoints every ckpt steps),
    checked against central differences (eps fd_eps) along fd_dirs seeded
    Gaussian directions.  Every time is device time between HIP events on
    the solver's stream: the plain forward steps (a pair of events around
    the steps), the recorded forward with the objective and the reversal
    (events inside astr_adjoint_enstrophy).  The cost ratio is the whole
    gradient (recorded forward + reversal) over the plain forward."""
    import torch

    import astr_gpu_amd as ag
    from astr_gpu_amd.autograd import velocity_state

    dt = 0.002
    p = ag.default_params(n, mach=0.3, reynolds=173.0)
    s = ag.AstrSolver(p, device=0)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.hit_state(p, seed)
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
    runs = 3
    fwd = []
    for _ in range(runs):
        upload(st)
        s.set_timing(2)
        s.steploop(nsteps, dt)
        fwd.append(float(s.timing()[7]))
    s.set_timing(0)
    t_fwd = float(np.median(fwd))

    def gradient(k):
        res = []
        for _ in range(runs):
            upload(st)
            J, qb, pb, ms = s.adjoint_enstrophy(nsteps, dt, k)
            res.append((float(ms[0]), float(ms[1]), J, qb, pb))
        rec = float(np.median([r[0] for r in res]))
        rev = float(np.median([r[1] for r in res]))
        return rec, rev, res[-1][2:]

    rec_all, rev_all, _ = gradient(nsteps)
    t_rec, t_rev, (J, qb, pb) = gradient(ckpt)
    # dJ/du0 through the state construction
    x = torch.tensor(u0, requires_grad=True)
    q0, pr0 = velocity_state(p, x)
    torch.autograd.backward([q0, pr0], [torch.from_numpy(qb), torch.from_numpy(pb)])
    g = x.grad.numpy()
    rng = np.random.default_rng(seed)
    errs = []
    for _ in range(fd_dirs):
        v = rng.standard_normal(u0.shape)
        ad = float(np.sum(g * v))
        fd = (objective(u0 + fd_eps * v) - objective(u0 - fd_eps * v)) / (2 * fd_eps)
        errs.append(abs(fd - ad) / abs(ad))
    s.close()
    return {
        "workload": f"discrete adjoint, HIT {n}^3 Mt 0.3 Re 173 seed {seed}, {nsteps} RK3 steps "
                    f"dt {dt}, checkpoints every {ckpt} steps, J = enstrophy of the final state",
        "gradient_over_forward": (t_rec + t_rev) / t_fwd,
        "reversal_over_forward": t_rev / t_fwd,
        "forward_ms": t_fwd, "recorded_forward_ms": t_rec, "reversal_ms": t_rev,
        "gradient_over_forward_all_stages_kept": (rec_all + rev_all) / t_fwd,
        "timing": f"device time between HIP events on the solver's stream, median of {runs} runs",
        "J": J, "fd_max_rel_err": max(errs, default=None), "fd_eps": fd_eps,
        "fd_directions": fd_dirs,
    }


def bench_adjoint(args):
    """--workload adjoint: the adjoint leg alone as the JSON line; the value
    is the whole gradient's cost over the plain forward's (lower is
    better)."""
    leg = adjoint_leg(args.n, args.nsteps, args.ckpt, args.seed, args.fd_dirs)
    out = {
        "metric": "adjoint/forward cost ratio (recorded forward + reversal over the plain "
                  "forward, device time)",
        "value": leg["gradient_over_forward"], "unit": "ratio", "n_gpus": 1,
        "steps": args.nsteps, "warmup": 1,
        "ms_per_step": leg["forward_ms"] / args.nsteps, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded decaying isotropic turbulence)",
        "config": {"workload": leg["workload"], "n": args.n, "parallelism": "single GPU"},
        "adjoint": leg, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def ensemble_leg(nbox=48, seed=1, steps=10, warmup=1, sizes=None):
    """48 independent periodic boxes of decaying isotropic turbulence (sizes
    drawn from {64,...,192} per axis, seeded), the central scheme with
    filter and viscous terms: aggregate grid-point-RK3-steps/s of the
    batched modes (a graph over the members' streams, direct launches on
    them, and the autotuned line-queue engine; the value is the fastest)
    against looping the plain solver over the boxes, each timed by the wall
    clock around its steps (all end synchronised)."""
    import astr_gpu_amd as ag

    dims = ag.ensemble_dims(nbox, seed, sizes or ag.solver.ENSEMBLE_SIZES)
    dt = 0.002
    # Re_lambda ~ 50 for the k^4 exp(-2 (k/4)^2) spectrum at sqrt(<u.u>) = 1:
    # lambda = 0.5, u' = 3^-1/2, so Re = 50 / (0.5 * 3^-1/2) = 173
    solvers = []
    for i, d in enumerate(dims):
        p = ag.default_params(d, mach=0.3, reynolds=173.0)
        s = ag.AstrSolver(p, device=0)
        s.upload_geom(*ag.box_metrics(p))
        st = ag.hit_state(p, ag.ensemble_member_seed(seed, i))
        s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        solvers.append(s)
    pts = float(sum(a * b * c for a, b, c in dims))
    # plain solver, box after box
    for s in solvers:
        s.steploop(warmup, dt)
    t0 = time.perf_counter()
    for s in solvers:
        s.steploop(steps, dt)
    el_loop = time.perf_counter() - t0
    # batched: one graph per step over all boxes, and the same launches made
    # directly on the members' streams
    b = ag.AstrBatch(solvers)
    b.rk3_steps(warmup, dt)
    t0 = time.perf_counter()
    b.rk3_steps(steps, dt)
    el_graph = time.perf_counter() - t0
    b.set_mode(1)
    b.rk3_steps(warmup, dt)
    t0 = time.perf_counter()
    b.rk3_steps(steps, dt)
    el_streams = time.perf_counter() - t0
    # the line-queue engine, after its autotuner's trial steps
    b.set_mode(2)
    t0 = time.perf_counter()
    tuning = b.autotune(dt)
    tune_s = time.perf_counter() - t0
    b.rk3_steps(warmup, dt)
    t0 = time.perf_counter()
    b.rk3_steps(steps, dt)
    el_queue = time.perf_counter() - t0
    b.close()
    el = min(el_graph, el_streams, el_queue)
    for s in solvers:
        s.close()
    value = pts * steps / el
    loop_value = pts * steps / el_loop
    return {
        "workload": f"{nbox} periodic boxes, sizes per axis from "
                    f"{list(sizes or ag.solver.ENSEMBLE_SIZES)} (seed {seed}), HIT Mt 0.3, "
                    f"Re 173, 643c, filter 0.49, RK3 dt {dt}",
        "boxes": [list(d) for d in dims], "points": pts, "steps": steps,
        "value": value, "loop_value": loop_value, "batched_over_loop": value / loop_value,
        "graph_value": pts * steps / el_graph, "streams_value": pts * steps / el_streams,
        "queue_value": pts * steps / el_queue,
        "ms_per_step": 1e3 * el / steps, "loop_ms_per_step": 1e3 * el_loop / steps,
        "queue_tuning": [{k: e[k] for k in ("kind", "L", "config", "meaning")} for e in tuning],
        "tune_s": tune_s,
    }


def bench_ensemble(args):
    """--workload ensemble: the ensemble leg alone as the JSON line."""
    leg = ensemble_leg(args.nbox, args.seed, args.steps, args.warmup)
    out = {
        "metric": "FP64 grid-point-RK3-steps/s, batched ensemble",
        "value": leg["value"], "unit": "grid-point-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": leg["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded decaying isotropic turbulence)",
        "config": {"workload": leg["workload"], "boxes": leg["boxes"], "points": leg["points"],
                   "parallelism": "single GPU"},
        "loop_value": leg["loop_value"], "batched_over_loop": leg["batched_over_loop"],
        "graph_value": leg["graph_value"], "streams_value": leg["streams_value"],
        "queue_value": leg["queue_value"], "queue_tuning": leg["queue_tuning"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def tgv_dt(n: int) -> float:
    """dt of the n^3 Taylor-Green box: 0.004 at 64^3 (the small-size
    workload), scaled with the spacing so that the acoustic CFL number
    (|u| + 1/Ma) dt / dx stays 0.45 (0.001 at 256^3, the reference deck's
    own dt; 0.004 at 256^3 would be 1.8, past the RK3 stability limit of
    the compact scheme)."""
    return 0.004 * 64 / n


TIMED_RUNS = 5  # the value is the median of this many timed runs


def tgv_leg(n: int, steps: int, warmup: int):
    """The headline: n^3 Taylor-Green RK3 steps with the statistics of every
    step (rkfirst's statcal, src/mainloop.F90:626), the box resident in HBM.
    Each of TIMED_RUNS runs of `steps` steps is bracketed by two HIP events
    on the solver's stream; the value is the median.  One more run with
    events between the kernel families gives the roofline's launch times."""
    import astr_gpu_amd as ag

    dt = tgv_dt(n)
    p = ag.default_params(n)
    s = ag.AstrSolver(p, device=0)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.tgv_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    del st
    s.steploop(warmup, dt, stats=True)
    runs_ms, walls = [], []
    stats = None
    for _ in range(TIMED_RUNS):
        s.set_timing(2)
        t0 = time.perf_counter()
        stats = s.steploop(steps, dt, stats=True)   # synchronizes before returning
        walls.append(time.perf_counter() - t0)
        runs_ms.append(float(s.timing()[7]))
    # kernel families timed with events between them (not the value: the
    # events wait at every stage)
    s.set_timing(1)
    s.steploop(steps, dt, stats=True)
    fam = s.timing()
    s.set_timing(0)
    copy_ms, copy_bytes = s.bench_copy(20)
    s.close()
    ms = float(np.median(runs_ms))
    return dict(n=n, dt=dt, steps=steps, ms=ms, runs_ms=runs_ms, walls=walls, fam=fam,
                copy_gbs=copy_bytes / (copy_ms * 1e-3) / 1e9,
                last_stats=[float(v) for v in np.asarray(stats)[-1]],
                finite=bool(np.all(np.isfinite(np.asarray(stats)))))


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
                    help="tgv: the headline (plain 256^3 steps, with the batched and the "
                         "adjoint legs unless --legs says otherwise); riemann2d: RK3 steps of "
                         "the 2-D Riemann problem (upwind, boundary conditions); ensemble, "
                         "adjoint: that leg alone")
    ap.add_argument("--legs", default="plain,batched,adjoint",
                    help="tgv workload: the parts of the metric to measure")
    ap.add_argument("--nsteps", type=int, default=40, help="adjoint: RK3 steps of the objective")
    ap.add_argument("--ckpt", type=int, default=8, help="adjoint: steps between kept states")
    ap.add_argument("--fd-dirs", type=int, default=8,
                    help="adjoint: directions of the central-difference check (0: none)")
    ap.add_argument("--adj-n", type=int, default=128, help="tgv workload: adjoint leg box size")
    ap.add_argument("--t-end", type=float, default=0.0,
                    help="riemann2d: also run the problem from t = 0 to this time (0.3 in "
                         "the benchmark) and report its wall time")
    ap.add_argument("--nbox", type=int, default=48)
    ap.add_argument("--ens-sizes", default=None,
                    help="tgv workload: axis sizes of the batched leg (default 64,...,192)")
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

    legs = set(args.legs.split(","))
    n = args.n
    r = tgv_leg(n, args.steps, args.warmup)
    fam = r["fam"]
    pts = float(n) ** 3
    ms_per_step = r["ms"] / args.steps
    value = pts / (ms_per_step * 1e-3)
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
        "metric": "FP64 grid-point-RK3-steps/s at 256^3 (plain vs batched); adjoint/fwd ratio",
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
        "data": "synthetic (Taylor-Green initial field; seeded turbulence for the batched and "
                "adjoint legs)",
        "config": {"workload": f"Taylor-Green vortex {n}^3 periodic, Re 1600, Ma 0.1, 643c, "
                               f"filter alpha 0.49, RK3 dt {r['dt']:g} (acoustic CFL 0.45), "
                               f"statistics every step", "n": n,
                   "parallelism": "single GPU"},
        "timing": {"method": "two HIP events on the solver's stream around each run of "
                             f"{args.steps} steps; value from the median of {TIMED_RUNS} runs",
                   "runs_ms": r["runs_ms"],
                   "wall_ms": [1e3 * w for w in r["walls"]]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, n),
                     "kernel": f"{FAMILY_KERNEL[dom]}, {FAMILY_LAUNCHES[dom]} launches/stage",
                     "algo_bytes_per_launch": algo, "avg_launch_ms": avg_s * 1e3},
        # the whole step against the reference's 9-pass byte model (4 224 B
        # per grid-point-RK3-step) and against the bytes of the kernels that
        # run here (ALGO_BYTES_PER_PT_STEP)
        "step_roofline": {
            "bytes_per_pt_step_reference_passes": REF_BYTES_PER_PT_STEP,
            "achieved_gbs": value * REF_BYTES_PER_PT_STEP / 1e9,
            "frac": value * REF_BYTES_PER_PT_STEP / (HBM_PEAK_GBS * 1e9),
            "bytes_per_pt_step_these_kernels": ALGO_BYTES_PER_PT_STEP,
            "frac_these_kernels": value * ALGO_BYTES_PER_PT_STEP / (HBM_PEAK_GBS * 1e9),
            "copy_gbs": r["copy_gbs"],
            "frac_of_copy": value * REF_BYTES_PER_PT_STEP / (r["copy_gbs"] * 1e9)},
        "family_ms_per_step": {k: v / args.steps for k, v in fam_ms.items()},
        "state": {"finite": r["finite"], "last_step_stats": r["last_stats"]},
    }
    if "batched" in legs:
        sizes = tuple(int(v) for v in args.ens_sizes.split(",")) if args.ens_sizes else None
        out["batched"] = ensemble_leg(args.nbox, args.seed, 10, 1, sizes)
        out["batched_over_loop"] = out["batched"]["batched_over_loop"]
    if "adjoint" in legs:
        out["adjoint"] = adjoint_leg(args.adj_n, args.nsteps, args.ckpt, args.seed, args.fd_dirs)
        out["adjoint_over_forward"] = out["adjoint"]["gradient_over_forward"]
        out["adjoint_fd_max_rel_err"] = out["adjoint"]["fd_max_rel_err"]
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(n, r["dt"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
