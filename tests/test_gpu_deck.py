"""Stand-alone decks on the GPU: the reference's example decks, shrunk,
run through astr_gpu_amd.deck, and their flowstate.dat checked against the
checker's statistics series (Taylor-Green) and against the state the steps
reach (Riemann, max_q)."""
import os

import numpy as np
import pytest

import astr_gpu_amd as ag
import pyoracle
from astr_gpu_amd import deck

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
HM = 5


def shrunk_case(tmp_path, name, inp, dims, maxstep, deltat):
    """The example deck with its grid size and controller replaced."""
    text = open(os.path.join(DECKS, name, "datin", inp)).read()
    old = {"tgv": "128,128,128", "riem2d": "400,400,0"}[name]
    assert old in text
    case = tmp_path / name
    (case / "datin").mkdir(parents=True)
    (case / "datin" / inp).write_text(text.replace(old, ",".join(str(v) for v in dims)))
    (case / "datin" / "controller").write_text(
        "#\n#\n#\n\n# lwsequ,lwslic,lavg,lcracon\nf, f, f, f\n\n"
        "# maxstep,feqchkpt,feqwsequ,feqslice,feqlist,feqavg\n"
        f"{maxstep}, 1000, 1000, 50, 1, 50\n\n# deltat\n{deltat!r}\n")
    return str(case)


def read_flowstate(path):
    with open(path) as f:
        head = f.readline().split()
        rows = [[float(v) for v in line.split()] for line in f if line.strip()]
    return head, np.array(rows)


def test_taylor_green_deck_against_the_checker(tmp_path):
    dims, maxstep, dt = (32, 32, 32), 4, 0.004
    print(f"case tgv deck {dims} maxstep {maxstep}", flush=True)
    case = shrunk_case(tmp_path, "tgv", "input.tgv", dims, maxstep, dt)
    path = deck.run_deck(case, log=None)
    head, rows = read_flowstate(path)
    assert head == ["nstep", "time", "kenergy", "enstophy", "dissipation"]
    assert rows.shape == (maxstep, 5)
    ob = pyoracle.OracleBox(dims, variant=0)
    ob.init_tgv()
    t = 0.0
    for k in range(maxstep + 1):
        st = ob.step(dt, stats=True)
        if k >= 1:
            row = rows[k - 1]
            assert row[0] == k
            assert abs(row[1] - t) <= 1e-13 * t
            for a, b in zip(row[2:], st):
                assert abs(a - b) <= 1e-10 * abs(b), (k, row, st)
        t += dt
    ob.close()


def test_hit_deck_against_the_checker(tmp_path):
    """The Taylor-Green deck with flowtype hit: the seeded turbulence field,
    stepped by the deck runner, against the checker stepped from the same
    field (its series of KE, enstrophy and dissipation)."""
    dims, maxstep, dt, seed = (32, 32, 32), 3, 0.002, 5
    print(f"case hit deck {dims} maxstep {maxstep} seed {seed}", flush=True)
    case = shrunk_case(tmp_path, "tgv", "input.tgv", dims, maxstep, dt)
    inp = os.path.join(case, "datin", "input.tgv")
    text = open(inp).read()
    assert "\ntgv\n" in text
    open(inp, "w").write(text.replace("\ntgv\n", "\nhit\n"))
    path = deck.run_deck(case, log=None, seed=seed)
    head, rows = read_flowstate(path)
    assert head == ["nstep", "time", "kenergy", "enstophy", "dissipation"]
    assert rows.shape == (maxstep, 5)
    d = deck.read_input(deck.find_input(case))
    assert d["flowtype"] == "hit"
    p = deck.params_from_deck(d)
    st = ag.hit_state(p, seed)
    ob = pyoracle.OracleBox(dims, variant=0, reynolds=p.reynolds, mach=p.mach,
                            prandtl=p.prandtl, gamma=p.gamma, ref_tem=p.ref_tem,
                            alfa=p.alfa_filter)
    for m in range(5):
        ob.set(f"q{m + 1}", ag.interior(st["q"][..., m], p))
    ob.set("rho", ag.interior(st["rho"], p))
    for c, nm in enumerate("uvw"):
        ob.set(nm, ag.interior(st["vel"][..., c], p))
    ob.set("prs", ag.interior(st["prs"], p))
    ob.set("tmp", ag.interior(st["tmp"], p))
    t = 0.0
    for k in range(maxstep + 1):
        ref = ob.step(dt, stats=True)
        if k >= 1:
            row = rows[k - 1]
            assert row[0] == k and abs(row[1] - t) <= 1e-13 * t
            for a, b in zip(row[2:], ref):
                assert abs(a - b) <= 1e-10 * abs(b), (k, row, ref)
        t += dt
    ob.close()


def test_riemann_deck_writes_max_q_of_each_stage_1_state(tmp_path):
    dims, maxstep, dt = (64, 64, 0), 3, 1e-4
    print(f"case riem2d deck {dims} maxstep {maxstep}", flush=True)
    case = shrunk_case(tmp_path, "riem2d", "input.riemann2d", dims, maxstep, dt)
    path = deck.run_deck(case, log=None)
    head, rows = read_flowstate(path)
    assert head == ["nstep", "time", "maxq1", "maxq2", "maxq3", "maxq4", "maxq5"]
    assert rows.shape == (maxstep, 7)
    # the same steps through the solver's own calls
    d = deck.read_input(deck.find_input(case))
    p = deck.params_from_deck(d)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p, lengths=(1.0, 1.0, 1.0)))
    st = ag.riemann2d_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.set_bc(tuple(d["bctype"]), ag.UDF_RIEMANN2D)
    s.time = 0.0
    for k in range(maxstep + 1):
        t = s.time
        s.boucon()
        s.qswap()
        if k >= 1:
            q = s.download_state()["q"]
            interior = q[HM:HM + dims[0] + 1, HM:HM + dims[1] + 1, HM:HM + 1, :]
            ref = interior.reshape(-1, 5).max(axis=0)
            row = rows[k - 1]
            assert row[0] == k and abs(row[1] - t) <= 1e-13 * t
            assert np.allclose(row[2:], ref, rtol=1e-12, atol=0), (k, row, ref)
        s.time_integration_rk(dt)
    s.close()


def test_qmax_is_the_largest_value_on_the_nodes():
    dims = (40, 36, 33)
    p = ag.default_params(dims)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    rng = np.random.default_rng(2)
    shp = ag.haloed_shape(p)
    q = np.asfortranarray(rng.standard_normal(shp + (5,)))
    # halo cells hold larger values, which max_q must not see
    q[:HM] += 10.0
    s.upload_state(q, np.asfortranarray(q[..., 0].copy()), np.zeros(shp + (3,), order="F"),
                   np.ones(shp, order="F"), np.ones(shp, order="F"))
    got = s.qmax()
    ref = q[HM:HM + dims[0] + 1, HM:HM + dims[1] + 1, HM:HM + dims[2] + 1].reshape(-1, 5).max(axis=0)
    assert list(got) == list(ref)
    s.close()
