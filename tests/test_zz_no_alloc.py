"""astr_stats_device, the entry point that allocates nothing (its partial
sums live in the library's fixed device array): through the C ABI, into a
device buffer, against astr_stats on the same state, on two calls in a
row, the second on a smaller box; both statistics within 1e-14 relative
(both sum in double-double, in different groupings)."""
import ctypes

import numpy as np
import pytest

import astr_gpu_amd as ag
from astr_gpu_amd import _lib

pytestmark = pytest.mark.gpu


def box(n, seed):
    p = ag.default_params(n, mach=0.3, reynolds=173.0)
    s = ag.AstrSolver(p)
    s.upload_geom(*ag.box_metrics(p))
    st = ag.hit_state(p, seed)
    s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
    s.steploop(2, 0.002)
    s.qswap()
    s.rhscal()
    return s


def test_stats_device_matches_stats_on_two_calls_the_second_smaller():
    import torch

    lib = _lib.load()
    out = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    for n, seed in ((48, 3), (32, 4)):
        print(f"case stats_device n={n}", flush=True)
        s = box(n, seed)
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        _lib.check(lib.astr_stats_device(s.ctx, ctypes.c_void_p(out.data_ptr())),
                   "astr_stats_device")
        s.synchronize()
        got = out.cpu().numpy()
        want = np.array(s.statcal())
        print(f"  device {got} host {want}", flush=True)
        assert np.all(np.isfinite(got))
        assert np.all(np.abs(got - want) <= 1e-14 * np.abs(want)), (got, want)
        s.close()


def test_stats_device_rejects_a_null_buffer():
    lib = _lib.load()
    s = box(16, 5)
    assert lib.astr_stats_device(s.ctx, None) == 1  # ASTR_E_ARG
    s.close()
