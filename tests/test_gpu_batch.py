"""Batched ensemble: one HIP graph per RK3 step over boxes of different
sizes.  Every graph replay must reproduce, bit for bit, the same steps taken
box by box with astr_rk3_step; the replays are checked one at a time, with
eager work and synchronisation in between."""
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


@pytest.mark.parametrize("convection,mode", [(0, 0), (1, 0), (0, 1)])
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
