"""Pins the pure-PyTorch FP64 unrolled FDTD (tests/torch_fdtd.py) to the C
oracle: the same objective J on the adjoint test case, to rounding; and its
autograd gradient to the oracle adjoint at seeded voxels."""
import numpy as np
import torch

from adjoint_case import objective, oracle_adjoint, seeded_voxels, small_case
from torch_fdtd import run


def test_torch_reference_objective_matches_oracle():
    cfg, d0, eps, mons, omega, steps = small_case(steps=160)
    J_t = float(run(cfg, torch.tensor(eps), d0, mons, omega, steps))
    J_o = objective(cfg, d0, eps, mons, omega, steps)
    assert abs(J_t / J_o - 1) < 1e-10


def test_torch_autograd_gradient_matches_oracle_adjoint():
    cfg, d0, eps, mons, omega, steps = small_case(steps=160)
    e = torch.tensor(eps, requires_grad=True)
    run(cfg, e, d0, mons, omega, steps).backward()
    g_eps = e.grad.numpy()                       # dJ/deps
    _, g_inv = oracle_adjoint(cfg, d0, eps, mons, omega, steps)
    g_from_adj = g_inv * (-1.0 / eps ** 2)       # dJ/d(1/eps) -> dJ/deps
    vox = seeded_voxels(eps.shape, n=16)
    a = np.array([g_eps[v] for v in vox])
    b = np.array([g_from_adj[v] for v in vox])
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-6
