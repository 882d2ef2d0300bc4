"""Batched ensemble: one HIP graph per RK3 step over boxes of different
sizes (mode 0: the members' own launches on their streams; mode 2: the
line-queue engine's launches over all members).  Every graph replay must
reproduce, bit for bit, the same steps taken box by box with
astr_rk3_step; the replays are checked one at a time, with eager work and
synchronisation in between."""
import numpy as np
import pytest

import astr_gpu_amd as ag

pytestmark = pytest.mark.gpu


def make_box(dims, seed, **over):
    p = ag.default_params(dims, mach=0.3, **over)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.hit_state(p, seed)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    return s


@pytest.mark.parametrize("convection,mode", [(0, 0), (1, 0), (0, 1), (0, 2)])
def test_batch_graph_replays_equal_box_by_box_steps(convection, mode):
    dims = [(32, 32, 32), (40, 36, 32), (48, 32, 40)]
    over = {} if convection == 0 else dict(convection=1, deriv_explicit=1, lfilter=1,
                                           diffterm=0)
    print(f"case batch convection={convection} mode={mode} dims={dims}", flush=True)
    eager = [make_box(d, 10 + i, **over) for i, d in enumerate(dims)]
    graph = [make_box(d, 10 + i, **over) for i, d in enumerate(dims)]
    b = ag.AstrBatch(graph)
    b.set_mode(mode)
    dt = 0.004
    for step in range(3):
        b.rk3_steps(1, dt)
        for s in eager:
            s.time_integration_rk(dt)
        for k, (se, sg) in enumerate(zip(eager, graph)):
            a, g = se.download_state(), sg.download_state()
            for nm in a:
                assert np.array_equal(a[nm], g[nm]), (step, k, nm)
        # eager work on a member between replays, then back to the graph
        st = graph[0].statcal()
        assert np.all(np.isfinite(st))
    # several replays in one call
    b.rk3_steps(2, dt)
    for s in eager:
        s.steploop(2, dt)
    for se, sg in zip(eager, graph):
        a, g = se.download_state(), sg.download_state()
        for nm in a:
            assert np.array_equal(a[nm], g[nm]), nm
    b.close()
    for s in eager + graph:
        s.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_batch_members_stepped_alone_in_between(mode):
    """Whole steps of one box alternate its q and qsave buffers (stage 1
    writes its result into the other one); a batch holding the box rebuilds
    its launches after them.  Plain steps on a member between batched steps
    (an odd number, so the buffers have swapped) must leave every member
    bit-equal to boxes stepped alone throughout."""
    dims = [(32, 32, 32), (40, 36, 32)]
    print(f"case batch interleaved mode={mode} dims={dims}", flush=True)
    eager = [make_box(d, 20 + i) for i, d in enumerate(dims)]
    mixed = [make_box(d, 20 + i) for i, d in enumerate(dims)]
    b = ag.AstrBatch(mixed)
    b.set_mode(mode)
    dt = 0.004
    for rnd in range(2):
        b.rk3_steps(1, dt)
        mixed[0].steploop(1, dt)
        mixed[1].steploop(3, dt)
        b.rk3_steps(1, dt)
        eager[0].steploop(3, dt)
        eager[1].steploop(5, dt)
        for k, (se, sm) in enumerate(zip(eager, mixed)):
            a, g = se.download_state(), sm.download_state()
            for nm in a:
                assert np.array_equal(a[nm], g[nm]), (rnd, k, nm)
    b.close()
    for s in eager + mixed:
        s.close()


def test_batch_rejects_mixed_or_unready_members():
    s1 = make_box((32, 32, 32), 1)
    p = ag.default_params((32, 32, 32))
    s2 = ag.AstrSolver(p)  # no state uploaded
    with pytest.raises(ag.AstrError):
        ag.AstrBatch([s1, s2])
    with pytest.raises(ag.AstrError):
        ag.AstrBatch([s1, s1])
    s1.close()
    s2.close()


def test_queue_engine_autotune_keeps_the_state_and_steps_bit_for_bit():
    """Boxes whose axes fall in three rows-per-lane classes (L 2, 3, 4): the
    autotuner's trial steps leave every member's state as it was; after
    tuning, the queue engine's steps equal box-by-box steps bit for bit, for
    the tuned choices and for each fixed configuration of the launches."""
    dims = [(40, 64, 96), (96, 40, 64), (64, 96, 40), (33, 33, 33)]
    print(f"case queue autotune dims={dims}", flush=True)
    eager = [make_box(d, 20 + i) for i, d in enumerate(dims)]
    batch = [make_box(d, 20 + i) for i, d in enumerate(dims)]
    b = ag.AstrBatch(batch)
    b.set_mode(2)
    dt = 0.004
    before = [s.download_state() for s in batch]
    rec = b.autotune(dt)
    after = [s.download_state() for s in batch]
    for x, y in zip(before, after):
        for nm in x:
            assert np.array_equal(x[nm], y[nm]), nm
    kinds = {(e["kind"], e["L"]) for e in rec}
    assert {(k, L) for k in ("filter", "gradient", "flux") for L in (2, 3, 4)} <= kinds
    for e in rec:
        assert 0 <= e["config"] < 8 and len(e["trial_us"]) == 8
        assert e["trial_us"][e["config"]] == min(e["trial_us"])
    print("  tuning:", [(e["kind"], e["L"], e["config"]) for e in rec], flush=True)
    for cfg in (None, 0, 7, 2, 5):
        if cfg is not None:
            b.set_tuning([{"kind": k, "L": L, "config": cfg} for k, L in kinds])
        b.rk3_steps(2, dt)
        for s in eager:
            s.steploop(2, dt)
        for k, (se, sg) in enumerate(zip(eager, batch)):
            x, y = se.download_state(), sg.download_state()
            for nm in x:
                assert np.array_equal(x[nm], y[nm]), (cfg, k, nm)
            assert se.time == sg.time
    b.close()
    for s in eager + batch:
        s.close()


def test_queue_engine_rejects_upwind_members_and_rechecks_members():
    s1 = make_box((32, 32, 32), 1, convection=1, deriv_explicit=1, lfilter=1, diffterm=0)
    s2 = make_box((32, 32, 32), 2)
    b = ag.AstrBatch([s1, s2])
    with pytest.raises(ag.AstrError, match="UNSUPPORTED"):
        b.set_mode(2)
    b.close()
    b = ag.AstrBatch([s2])
    b.set_mode(2)
    b.rk3_steps(1, 0.004)
    # a member given a user boundary condition after the batch was made is
    # refused at the next batched step
    s2.set_bc((0, 1, 1, 1, 1, 1), ag.UDF_RIEMANN2D)
    with pytest.raises(ag.AstrError, match="UNSUPPORTED"):
        b.rk3_steps(1, 0.004)
    b.close()
    s1.close()
    s2.close()
