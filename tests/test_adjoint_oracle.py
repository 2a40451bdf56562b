"""Pins the adjoint algorithm of dtm_adjoint_design_gradient on the FP64
oracle: the transposed leapfrog step run as the forward step on
(lambda_E/eps, -lambda_H) gives dJ/d(1/eps) equal to central finite
differences of J (CPML on every face, soft source, random design box)."""
import numpy as np

from adjoint_case import fd_gradient, objective, oracle_adjoint, seeded_voxels, small_case


def test_oracle_adjoint_matches_finite_differences():
    cfg, d0, eps, mons, omega, steps = small_case()
    J, grad = oracle_adjoint(cfg, d0, eps, mons, omega, steps)
    assert abs(J / objective(cfg, d0, eps, mons, omega, steps) - 1) < 1e-12
    vox = seeded_voxels(eps.shape, n=6)
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    ga = np.array([grad[v] for v in vox])
    assert np.all(np.abs(gfd) > 0)
    assert np.linalg.norm(ga - gfd) / np.linalg.norm(gfd) < 1e-5


def test_oracle_adjoint_sign_error_would_fail():
    """The check is sensitive: dropping the eps^2 of the gradient sum (the
    chain from lambda_E = eps zeta and dt curl H = eps dE, a plausible
    slip) moves the gradient far beyond the tolerance."""
    cfg, d0, eps, mons, omega, steps = small_case(steps=200)
    J, grad = oracle_adjoint(cfg, d0, eps, mons, omega, steps)
    vox = seeded_voxels(eps.shape, n=3)
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    wrong = np.array([grad[v] / eps[v] ** 2 for v in vox])  # eps^2 factor dropped
    assert np.linalg.norm(wrong - gfd) / np.linalg.norm(gfd) > 1e-2


def test_oracle_adjoint_exact_beside_the_cpml():
    """Design box and monitors in the cells right next to the CPML slabs
    (x: the design box starts at the first cell after the 6-cell layer, the
    monitors sit in the last cell before the far layer): the adjoint, which
    steps the adjoint fields with the forward CPML update, still matches
    central differences (measured 6e-7).  The forward CPML update differs
    from the exact transpose only inside the layer, so what reaches cells
    outside it differs at the layer's reflection level.  (With the design
    box two cells into the layer and the monitors inside it the same
    comparison gives 7.5e-3: dtm_adjoint_design_gradient refuses that.)"""
    cfg, d0, eps, mons, omega, steps = small_case(steps=220)
    d0 = (cfg.npml[0], 10, 11)
    mons = [(cfg.nx - cfg.npml[1] - 1, 12 + a, 12 + b) for a in range(4) for b in range(4)]
    cfg.soft[0].i = 18
    J, grad = oracle_adjoint(cfg, d0, eps, mons, omega, steps)
    vox = [(0, 3, 3), (1, 4, 2), (5, 5, 3)]
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    ga = np.array([grad[v] for v in vox])
    assert np.all(np.abs(gfd) > 0)
    assert np.linalg.norm(ga - gfd) / np.linalg.norm(gfd) < 1e-5
