"""Checkpoint and restart on the GPU: a run that stops, writes its
checkpoint, and is resumed by a new solver from the files reaches the state
of the run that never stopped, bit for bit; and a deck restarted with
lrestart continues its flowstate.dat with the lines of the straight run."""
import os

import numpy as np
import pytest

import astr_gpu_amd as ag
from astr_gpu_amd import checkpoint as ck
from astr_gpu_amd import deck
from test_gpu_deck import read_flowstate, shrunk_case

pytestmark = pytest.mark.gpu


def tgv_solver(n):
    p = ag.default_params(n)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.tgv_state(p)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    return s


def test_resume_from_a_checkpoint_is_bit_for_bit(tmp_path):
    print("case checkpoint resume 32^3", flush=True)
    dt = 0.004
    a = tgv_solver(32)
    a.steploop(6, dt)
    want = a.download_state()
    t_want = a.time
    a.close()
    b = tgv_solver(32)
    b.steploop(3, dt)
    ck.write_checkpoint(b, str(tmp_path), 3)
    b.close()
    c = ag.AstrSolver(ag.default_params(32))
    c.upload_geom(*ag.box_metrics(c.p))
    assert ck.read_checkpoint(c, str(tmp_path)) == 3
    c.steploop(3, dt)
    got = c.download_state()
    assert c.time == t_want
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    c.close()


def test_deck_restart_continues_the_series(tmp_path):
    """maxstep 8 straight, against maxstep 8 restarted from the checkpoint
    the first run wrote at nstep 4 (feqchkpt 4): the restarted flowstate.dat
    keeps lines 1..3, and its lines 5..8 equal the straight run's (line 4,
    rkfirst's first call after the restart, is not written, as in the
    reference)."""
    dims, maxstep, dt = (24, 24, 24), 8, 0.004
    print(f"case deck restart {dims}", flush=True)
    straight = shrunk_case(tmp_path / "a", "tgv", "input.tgv", dims, maxstep, dt)
    _, rows_a = read_flowstate(deck.run_deck(straight, log=None))
    case = shrunk_case(tmp_path / "b", "tgv", "input.tgv", dims, maxstep, dt)
    ctl = os.path.join(case, "datin", "controller")
    text = open(ctl).read().replace(f"{maxstep}, 1000,", f"{maxstep}, 4,")
    open(ctl, "w").write(text)
    # first leg stops after nstep 5 (the checkpoint at the start of step 4)
    deck.run_deck(case, maxstep=5, log=None)
    assert os.path.exists(os.path.join(case, "outdat", "flowfield.npz"))
    inp = os.path.join(case, "datin", "input.tgv")
    src = open(inp).read()
    lines = src.split("\n")
    k = next(i for i, ln in enumerate(lines) if ln.startswith("# lrestar"))
    lines[k + 1] = "t"
    open(inp, "w").write("\n".join(lines))
    _, rows_b = read_flowstate(deck.run_deck(case, log=None))
    assert [int(r[0]) for r in rows_b] == [1, 2, 3, 5, 6, 7, 8]
    assert np.array_equal(rows_b[:3], rows_a[:3])
    assert np.array_equal(rows_b[3:], rows_a[4:])
