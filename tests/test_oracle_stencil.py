"""Pins the oracle's (2,4) Yee update (genAlg.py:182-301) with facts the
mathematics fixes, chosen so that a dropped term, a wrong sign, a wrong
stencil weight or a transposed index fails at least one of them:

* a periodic plane-wave eigenmode advances with the closed-form discrete
  frequency sin(w dt/2) = (dt/2)|K|, K_a = (2/d)[(9/8) sin(k d/2) - (1/24) sin(3k d/2)];
* the leapfrog invariant sum(eps E^n.E^n) + H^{n-1/2}.H^{n+1/2} is conserved
  exactly in a closed (zero-field or periodic) box -- only true when the H
  stencil is the exact negative transpose of the E stencil;
* the discrete divergence of E stays zero in vacuum (the (2,4) differences
  along different axes commute);
* the stability edge of the (2,4) scheme is Courant 6/7 (DTmxw.cu:332).
"""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs
from oracle.oracle import Oracle


def periodic_cfg(n, d, courant):
    cfg = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, courant))
    cfg.bc[:] = [abi.BC_PERIODIC] * 3
    return cfg


@pytest.mark.parametrize("axis,pol", [(0, 2), (1, 0), (2, 1), (0, 1)])
def test_axis_eigenmode_closed_form(axis, pol):
    n = [6, 6, 6]
    n[axis] = 36
    d = 0.1
    cfg = periodic_cfg(n, d, 0.7)
    o = Oracle(cfg)
    m = 5
    k = 2 * np.pi * m / (n[axis] * d)
    K = (2 / d) * (9 / 8 * np.sin(k * d / 2) - 1 / 24 * np.sin(3 * k * d / 2))
    w = 2 / cfg.dt * np.arcsin(cfg.dt * K / 2)
    wt = 2 / cfg.dt * np.sin(w * cfg.dt / 2)
    kv = np.zeros(3)
    kv[axis] = K
    e = np.zeros(3)
    e[pol] = 1.0
    hdir = np.cross(kv, e) / wt  # H = K x E / w~ for e^{i(k.r - wt)}
    shape = tuple(n)
    for c in range(3):
        pos = configs.component_positions(shape, c, d, d, d)
        o.set_field(c, e[c] * np.cos(k * pos[axis]) * np.ones(shape))
        pos = configs.component_positions(shape, 3 + c, d, d, d)
        o.set_field(3 + c, hdir[c] * np.cos(k * pos[axis] + w * cfg.dt / 2) * np.ones(shape))
    N = 200
    o.step(N)
    pos = configs.component_positions(shape, pol, d, d, d)
    ref = np.cos(k * pos[axis] - w * N * cfg.dt) * np.ones(shape)
    assert np.abs(o.field(pol) - ref).max() < 1e-11
    # the continuous frequency would be visibly wrong: the check is sensitive
    wrong = np.cos(k * pos[axis] - k * N * cfg.dt) * np.ones(shape)
    assert np.abs(o.field(pol) - wrong).max() > 1e-3


def test_oblique_eigenmode_closed_form():
    n = (24, 16, 12)
    d = 0.1
    cfg = periodic_cfg(n, d, 0.8)
    o = Oracle(cfg)
    mvec = np.array([3, 2, 1])
    k = 2 * np.pi * mvec / (np.array(n) * d)
    K = (2 / d) * (9 / 8 * np.sin(k * d / 2) - 1 / 24 * np.sin(3 * k * d / 2))
    w = 2 / cfg.dt * np.arcsin(cfg.dt * np.linalg.norm(K) / 2)
    wt = 2 / cfg.dt * np.sin(w * cfg.dt / 2)
    e = np.cross(K, [0.3, -0.2, 1.0])
    e /= np.linalg.norm(e)          # E perpendicular to the discrete K
    h = np.cross(K, e) / wt
    for c in range(3):
        x, y, z = configs.component_positions(n, c, d, d, d)
        o.set_field(c, e[c] * np.cos(k[0] * x + k[1] * y + k[2] * z))
        x, y, z = configs.component_positions(n, 3 + c, d, d, d)
        o.set_field(3 + c, h[c] * np.cos(k[0] * x + k[1] * y + k[2] * z + w * cfg.dt / 2))
    N = 150
    o.step(N)
    for c in range(3):
        x, y, z = configs.component_positions(n, c, d, d, d)
        ref = e[c] * np.cos(k[0] * x + k[1] * y + k[2] * z - w * N * cfg.dt)
        assert np.abs(o.field(c) - ref).max() < 1e-10


def leapfrog_invariant(o, eps_fields):
    """W_m = sum eps E^m.E^m + H^{m-1/2}.H^{m+1/2}: snapshot (E^m, H^{m-1/2})
    after m steps, advance one step for H^{m+1/2}."""
    E = [o.field(c) for c in range(3)]
    H0 = [o.field(3 + c) for c in range(3)]
    o.step(1)
    H1 = [o.field(3 + c) for c in range(3)]
    return sum(float((eps_fields[c] * E[c] ** 2).sum()) + float((H0[c] * H1[c]).sum()) for c in range(3))


@pytest.mark.parametrize("bc", [abi.BC_ZERO, abi.BC_PERIODIC])
def test_leapfrog_energy_invariant(bc):
    n = (14, 11, 9)
    d = 0.05
    cfg = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, 0.8))
    cfg.bc[:] = [bc] * 3
    cfg.nmat = 2
    cfg.mat[1].eps_inf = 3.0
    rng = np.random.RandomState(3)
    mats = [rng.randint(0, 2, size=n).astype(np.uint8) for _ in range(3)]
    o = Oracle(cfg)
    o.set_materials(*mats)
    for c in range(6):
        o.set_field(c, rng.standard_normal(n))
    eps = [np.where(m == 1, 3.0, 1.0) for m in mats]
    w0 = leapfrog_invariant(o, eps)
    for _ in range(5):
        o.step(40)
        w = leapfrog_invariant(o, eps)
        assert abs(w - w0) < 1e-11 * abs(w0)


def test_leapfrog_invariant_sees_loss():
    """The invariant as a detector: a conductive medium is not conservative."""
    n = (10, 9, 8)
    d = 0.05
    cfg = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, 0.8))
    cfg.mat[0].sigma = 0.5
    o = Oracle(cfg)
    rng = np.random.RandomState(4)
    for c in range(6):
        o.set_field(c, rng.standard_normal(n))
    w0 = leapfrog_invariant(o, [np.ones(n)] * 3)
    o.step(30)
    w1 = leapfrog_invariant(o, [np.ones(n)] * 3)
    assert w1 < 0.99 * w0


def fwd(F, axis):
    """(2,4) forward difference (E located at the lower half-point), zero outside."""
    pad = [(0, 0)] * 3
    pad[axis] = (1, 2)
    P = np.pad(F, pad)
    sl = lambda a, b: tuple(slice(a, P.shape[x] - b) if x == axis else slice(None) for x in range(3))
    fm1, f0, fp1, fp2 = P[sl(0, 3)], P[sl(1, 2)], P[sl(2, 1)], P[sl(3, 0)]
    return 27 * (fp1 - f0) - (fp2 - fm1)


def test_div_e_stays_zero_in_vacuum():
    n = (16, 14, 12)
    d = 0.05
    cfg = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, 0.8))
    o = Oracle(cfg)
    rng = np.random.RandomState(5)
    for c in range(3):
        o.set_field(3 + c, rng.standard_normal(n))
    o.step(25)
    E = [o.field(c) for c in range(3)]
    # div E at the all-odd node (2i+1,2j+1,2k+1): Ex sits at even x, etc.
    div = fwd(E[0], 0) + fwd(E[1], 1) + fwd(E[2], 2)
    scale = max(np.abs(fwd(E[0], 0)).max(), 1e-30)
    assert np.abs(div).max() < 1e-12 * scale * 1e3
    assert scale > 1e-3


@pytest.mark.parametrize("courant,stable", [(6 / 7 * 0.995, True), (6 / 7 * 1.02, False)])
def test_stability_edge_is_six_sevenths(courant, stable):
    n = (12, 12, 12)
    d = 0.05
    cfg = periodic_cfg(n, d, courant)
    o = Oracle(cfg)
    rng = np.random.RandomState(6)
    for c in range(6):
        o.set_field(c, rng.standard_normal(n))
    e0 = sum(float((o.field(c) ** 2).sum()) for c in range(6))
    o.step(600)
    e1 = sum(float((o.field(c) ** 2).sum()) for c in range(6))
    if stable:
        assert e1 < 10 * e0
    else:
        assert e1 > 1e6 * e0
