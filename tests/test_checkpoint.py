"""Checkpoint files on the CPU (astr_gpu_amd.checkpoint): the /restart/
namelist of auxiliary.txt, the flow-field archive's variables, the
consistency check of readcheckpoint, the reference's restart path (q from
the primitives), and flowstate.dat resumed as listinit does.  A stand-in
for the solver keeps the state in host arrays."""
import numpy as np
import pytest

import astr_gpu_amd as ag
from astr_gpu_amd import checkpoint as ck
from astr_gpu_amd import deck

HM = 5


class HostBox:
    """download_state / upload_state / time of a solver, in host memory."""

    def __init__(self, p, st=None):
        self.p = p
        self.st = st
        self.time = 0.0

    def download_state(self):
        return {k: v.copy(order="F") for k, v in self.st.items()}

    def upload_state(self, q, rho, vel, prs, tmp):
        self.st = {"q": q, "rho": rho, "vel": vel, "prs": prs, "tmp": tmp}


def random_state(p, seed):
    rng = np.random.default_rng(seed)
    s = ag.haloed_shape(p)
    return {"q": np.asfortranarray(rng.standard_normal(s + (5,))),
            "rho": np.asfortranarray(1.0 + 0.1 * rng.standard_normal(s)),
            "vel": np.asfortranarray(rng.standard_normal(s + (3,))),
            "prs": np.asfortranarray(rng.standard_normal(s)),
            "tmp": np.asfortranarray(rng.standard_normal(s))}


def test_auxiliary_namelist_round_trip_and_reference_text(tmp_path):
    path = tmp_path / "auxiliary.txt"
    ck.write_auxiliary(str(path), 120, filenumb=3)
    assert ck.read_auxiliary(str(path)) == {"nstep": 120, "filenumb": 3, "fnumslic": 0,
                                            "ninflowslice": 0, "nsamples": 0}
    # what a Fortran namelist write of /restart/ looks like, any case, and
    # names missing from the record
    path.write_text("&RESTART\n NSTEP=40,\n FILENUMB=2, fnumslic=1 /\n")
    assert ck.read_auxiliary(str(path))["nstep"] == 40
    assert ck.read_auxiliary(str(path))["fnumslic"] == 1
    assert ck.read_auxiliary(str(path))["nsamples"] == 0


def test_checkpoint_restores_the_state_bit_for_bit(tmp_path):
    p = ag.default_params((12, 10, 8))
    a = HostBox(p, random_state(p, 1))
    a.time = 0.123456789
    ck.write_checkpoint(a, str(tmp_path), 77)
    b = HostBox(p, random_state(p, 2))
    assert ck.read_checkpoint(b, str(tmp_path)) == 77
    assert b.time == a.time
    for k in a.st:
        assert np.array_equal(a.st[k], b.st[k]), k
    # the reference's variables, on the nodes 0..n
    with np.load(tmp_path / "flowfield.npz") as z:
        assert int(z["nstep"]) == 77 and float(z["time"]) == a.time
        I = (slice(HM, HM + 13), slice(HM, HM + 11), slice(HM, HM + 9))
        assert np.array_equal(z["ro"], a.st["rho"][I])
        assert np.array_equal(z["u3"], a.st["vel"][I + (2,)])
        assert np.array_equal(z["t"], a.st["tmp"][I])


def test_restart_stops_on_inconsistent_files(tmp_path):
    p = ag.default_params(8)
    a = HostBox(p, random_state(p, 3))
    ck.write_checkpoint(a, str(tmp_path), 10)
    ck.write_auxiliary(str(tmp_path / "auxiliary.txt"), 11)
    with pytest.raises(ck.CheckpointError, match="not consistent"):
        ck.read_checkpoint(HostBox(p), str(tmp_path))
    ck.write_auxiliary(str(tmp_path / "auxiliary.txt"), 10)
    with pytest.raises(ck.CheckpointError, match="box"):
        ck.read_checkpoint(HostBox(ag.default_params(10)), str(tmp_path))


def test_reference_variables_only_rebuild_q_from_the_primitives(tmp_path):
    """An archive with only ro, u1..u3, p, t (what the reference writes):
    updateq's q = (rho, rho u, rho v, rho w, rho (T const1 + |u|^2 / 2))."""
    p = ag.default_params(8, mach=0.3)
    n = 9
    rng = np.random.default_rng(4)
    ro = 1.0 + 0.1 * rng.random((n, n, n))
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    t = 1.0 + 0.1 * rng.random((n, n, n))
    pr = ro * t / (p.gamma * p.mach ** 2)
    np.savez(tmp_path / "flowfield.npz", nstep=np.int64(5), time=np.float64(0.5), ro=ro,
             u1=u[0], u2=u[1], u3=u[2], p=pr, t=t, dims=np.array([8, 8, 8]))
    ck.write_auxiliary(str(tmp_path / "auxiliary.txt"), 5)
    b = HostBox(p)
    assert ck.read_checkpoint(b, str(tmp_path)) == 5 and b.time == 0.5
    I = (slice(HM, HM + n),) * 3
    q = b.st["q"]
    const1 = 1.0 / (p.gamma * (p.gamma - 1.0) * p.mach ** 2)
    assert np.allclose(q[I + (0,)], ro, rtol=0, atol=0)
    assert np.allclose(q[I + (2,)], ro * u[1], rtol=1e-15)
    e = ro * (t * const1 + 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2))
    assert np.allclose(q[I + (4,)], e, rtol=1e-15)
    assert not np.any(q[:HM])


def test_flowstate_resumed_before_the_restart_step(tmp_path):
    path = tmp_path / "flowstate.dat"
    lines = [deck.flowstate_header(["nstep", "time", "a"]) + "\n"]
    lines += [deck.flowstate_line(k, 0.1 * k, [float(k)]) + "\n" for k in range(1, 9)]
    path.write_text("".join(lines))
    kept = deck.resumed_lines(str(path), 5)
    assert kept == lines[:5]  # header and nstep 1..4
    assert deck.resumed_lines(str(tmp_path / "none.dat"), 5) is None
