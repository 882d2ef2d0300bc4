"""The PyTorch-CPU forward (the autograd reference of the adjoint) against
the checker's main-path steps, and its gradient against central
differences."""
import numpy as np
import pytest
import torch

import astr_gpu_amd as ag
import pyoracle as po
from torch_forward import HM, Box


def hit_u0(dims, seed, mach=0.3):
    p = ag.default_params(dims, mach=mach, reynolds=173.0)
    st = ag.hit_state(p, seed)
    u = ag.interior(st["vel"], p)[:-1, :-1, :-1]
    return p, np.moveaxis(u, -1, 0).copy()


def test_torch_forward_matches_the_checker_steps():
    dims = (16, 18, 20)
    p, u0 = hit_u0(dims, 3)
    box = Box(dims, reynolds=p.reynolds, mach=p.mach)
    dt = 0.004
    with torch.no_grad():
        Q, P = box.initial(torch.tensor(u0))
        for _ in range(2):
            Q, P = box.step(Q, P, dt)
    ob = po.OracleBox(dims, variant=0, reynolds=p.reynolds, mach=p.mach)
    Q0, P0 = box.initial(torch.tensor(u0))
    I = box.I
    for m in range(5):
        ob.set(f"q{m + 1}", Q0[m][I].numpy())
    for k, name in enumerate(("rho", "u", "v", "w", "prs", "tmp")):
        ob.set(name, P0[k][I].numpy())
    for _ in range(2):
        ob.step(dt)
    for m in range(5):
        ref = ob.get(f"q{m + 1}")
        got = Q[m][I].numpy()
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, (m, err)
    for k, name in enumerate(("rho", "u", "v", "w", "prs", "tmp")):
        ref = ob.get(name)
        got = P[k][I].numpy()
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, (name, err)
    ob.close()


def test_torch_gradient_against_central_differences():
    dims = (16, 16, 16)
    p, u0 = hit_u0(dims, 5)
    box = Box(dims, reynolds=p.reynolds, mach=p.mach)
    x = torch.tensor(u0, requires_grad=True)
    J = box.objective(x, 2, 0.004)
    J.backward()
    g = x.grad.numpy()
    rng = np.random.default_rng(0)
    eps = 1e-6
    for _ in range(2):
        v = rng.standard_normal(u0.shape)
        with torch.no_grad():
            jp = float(box.objective(torch.tensor(u0 + eps * v), 2, 0.004))
            jm = float(box.objective(torch.tensor(u0 - eps * v), 2, 0.004))
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(g * v))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
