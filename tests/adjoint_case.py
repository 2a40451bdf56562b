"""Shared small inverse-design case for the adjoint tests: soft Ez dipole,
random design box, a 4x4 monitor plane behind it, one frequency."""
import copy

import numpy as np

from dtmaxwell4_amd import abi, configs
from oracle.oracle import Oracle


def small_case(npml=6, steps=260, seed=3):
    nx, ny, nz, d = 36, 28, 28, 0.05
    cfg = abi.default_config(nx, ny, nz, d, d, d, configs.dt_for(d, d, d))
    cfg.npml[:] = [npml] * 6
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EZ, 10, 14, 14
    so.waveform, so.amp, so.w, so.tw, so.t0 = abi.WAVE_GAUSS_SINE, 1.0, 2 * np.pi / 0.5, 0.25, 0.8
    d0 = (14, 10, 11)
    eps = 1.0 + 3.0 * np.random.RandomState(seed).uniform(0, 1, size=(8, 8, 6))
    mons = [(25, 12 + a, 12 + b) for a in range(4) for b in range(4)]
    return cfg, d0, eps, mons, 2 * np.pi / 0.5, steps


def downsized_case(steps=800):
    """The inverse-design workload downsized to 48^3: CPML 8, design box
    16x16x8 of rho ~ U[0,1] (seeded, 3-cell box smoothing), eps = 1 + 11 rho,
    soft Ez dipole 4 cells before the box, 4x4 monitor plane 4 cells behind
    it, one frequency; the Gaussian pulse is shortened (width 1.5 periods,
    delay 6) so that 800 steps carry it through the box."""
    w = configs.inverse_design(nx=48, ny=48, nz=48, npml=8, design=(16, 16, 8), steps=steps, monitor=4)
    so = w.cfg.soft[0]
    f0 = w.notes["f0"]
    so.tw, so.t0 = 1.5 / f0, 6.0 / f0
    d0 = w.design[:3]
    eps = w.design[3]
    mons = [p[1:] for p in w.probes]
    return w.cfg, d0, eps, mons, float(w.omegas[0]), steps, w.notes["rho"]


def objective(cfg, d0, eps, mons, omega, steps):
    """J = sum_m |sum_n Ez_m^{n+1} e^{i w (n+1) dt} dt|^2 from the FP64 oracle."""
    o = Oracle(cfg)
    o.set_design_eps(*d0, eps)
    for (i, j, k) in mons:
        o.add_probe(abi.EZ, i, j, k)
    o.set_dft(np.array([omega]))
    o.step(steps)
    F = o.dft()[:, 0]
    o.close()
    return float((np.abs(F) ** 2).sum())


def fd_gradient(cfg, d0, eps, mons, omega, steps, voxels, rel=1e-3):
    """Central differences of J in 1/eps at the given design voxels."""
    out = []
    for v in voxels:
        ie0 = 1.0 / eps[v]
        h = rel * ie0
        ep, em = eps.copy(), eps.copy()
        ep[v], em[v] = 1.0 / (ie0 + h), 1.0 / (ie0 - h)
        out.append((objective(cfg, d0, ep, mons, omega, steps) - objective(cfg, d0, em, mons, omega, steps)) / (2 * h))
    return np.array(out)


def oracle_adjoint(cfg, d0, eps, mons, omega, steps):
    """The library's adjoint algorithm restated on the FP64 oracle: forward
    with the design-box E history, then the forward step applied backward to
    (zeta, eta) = (lambda_E / eps, -lambda_H) with dJ/dE injected at the
    monitors; dJ/d(1/eps) = sum_n eps^2 zeta^{n+1} . (E^{n+1} - E^n)."""
    i0, j0, k0 = d0
    ni, nj, nk = eps.shape
    sl = (slice(i0, i0 + ni), slice(j0, j0 + nj), slice(k0, k0 + nk))
    o = Oracle(cfg)
    o.set_design_eps(*d0, eps)
    hist = [np.stack([o.field(c)[sl] for c in range(3)])]
    ser = []
    for _ in range(steps):
        o.step(1)
        ez = o.field(abi.EZ)
        ser.append(np.array([ez[m] for m in mons]))
        hist.append(np.stack([o.field(c)[sl] for c in range(3)]))
    o.close()
    dt = cfg.dt
    t = (np.arange(steps) + 1) * dt
    F = (np.array(ser) * np.exp(1j * omega * t)[:, None] * dt).sum(0)
    J = float((np.abs(F) ** 2).sum())
    cfg2 = copy.deepcopy(cfg)
    cfg2.nsoft = 0
    a = Oracle(cfg2)
    a.set_design_eps(*d0, eps)

    def inject(n):  # monitors sit in vacuum (1/eps = 1)
        g = 2 * dt * (F.real * np.cos(omega * n * dt) + F.imag * np.sin(omega * n * dt))
        ez = a.field(abi.EZ)
        for q, m in enumerate(mons):
            ez[m] += g[q]
        a.set_field(abi.EZ, ez)

    inject(steps)
    grad = np.zeros(eps.shape)
    for n in range(steps - 1, -1, -1):
        z = np.stack([a.field(c)[sl] for c in range(3)])
        grad += eps ** 2 * (z * (hist[n + 1] - hist[n])).sum(0)
        a.step(1)
        if n >= 1:
            inject(n)
    a.close()
    return J, grad


def seeded_voxels(shape, n=16, seed=11):
    rng = np.random.RandomState(seed)
    out = []
    while len(out) < n:
        v = tuple(int(rng.randint(0, s)) for s in shape)
        if v not in out:
            out.append(v)
    return out
