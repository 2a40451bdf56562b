"""A pure-PyTorch FP64 unrolled FDTD of the same equations (test
infrastructure): (2,4) staggered update (genAlg.py:182-301), CPML with the
sigma grading of DTmxw.cu:220-249, zero padding outside the grid, soft
current sources and a design box of scalar permittivity.  Written
independently of the C oracle so that autograd through it is a second
reference for the adjoint gradient; its objective is pinned to the oracle's.

Cell (i,j,k) holds Ex(2i,2j+1,2k+1), Ey(2i+1,2j,2k+1), Ez(2i+1,2j+1,2k),
Hx(2i+1,2j,2k), Hy(2i,2j+1,2k), Hz(2i,2j,2k+1)."""
import math

import numpy as np
import torch
import torch.nn.functional as F

from dtmaxwell4_amd import purcell


def sigma(h, n, plo, phi, d, m, R, vmax):
    """CPML grading at half-coordinate h: sigma_max ((depth)/L)^m over each
    slab, sigma_max = vmax ln(1/R) (m+1) / (2 L)."""
    if plo > 0 and h < 2 * plo:
        L = plo * d
        return vmax * math.log(1 / R) * (m + 1) / (2 * L) * ((2 * plo - h) / (2 * plo)) ** m
    if phi > 0 and h > 2 * (n - phi):
        L = phi * d
        return vmax * math.log(1 / R) * (m + 1) / (2 * L) * ((h - 2 * (n - phi)) / (2 * phi)) ** m
    return 0.0


def _shift(t, axis, s):
    """t[..., i+s, ...] along axis with zeros outside the grid."""
    n = t.shape[axis]
    pad = [0, 0] * 3
    # F.pad orders pads from the last dimension backwards
    p = 2 * (2 - axis)
    if s > 0:
        pad[p + 1] = s
        tp = F.pad(t[None, None], pad[::1])[0, 0]
        return tp.narrow(axis, s, n)
    pad[p] = -s
    tp = F.pad(t[None, None], pad)[0, 0]
    return tp.narrow(axis, 0, n)


def d_fwd(f, axis, d):
    """Derivative at the lower neighbour's point: samples i-1, i, i+1, i+2 (E from H)."""
    return (27 * (_shift(f, axis, 1) - f) - (_shift(f, axis, 2) - _shift(f, axis, -1))) / (24 * d)


def d_bwd(f, axis, d):
    """Derivative from samples i-2, i-1, i, i+1 (H from E)."""
    return (27 * (f - _shift(f, axis, -1)) - (_shift(f, axis, 1) - _shift(f, axis, -2))) / (24 * d)


def run(cfg, eps_design, d0, monitors, omega, steps, device=None):
    """J = sum_m |sum_n Ez_m^{n+1} e^{i w (n+1) dt} dt|^2; eps_design is a
    torch float64 tensor (it may require grad).  `device` places the unrolled
    run (default: the CPU)."""
    dt = cfg.dt
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    n3 = (nx, ny, nz)
    dd = (cfg.dx, cfg.dy, cfg.dz)
    kw = dict(dtype=torch.float64, device=device)
    # CPML b, c along each axis at E (odd) and H (even) half-coordinates
    bc = {}
    for a in range(3):
        for kind, off in (("E", 1), ("H", 0)):
            s = np.array([sigma(2 * v + off, n3[a], cfg.npml[2 * a], cfg.npml[2 * a + 1], dd[a], cfg.pml_order,
                                cfg.pml_R, cfg.pml_vmax) for v in range(n3[a])])
            b = np.exp(-s * dt)
            shape = [1, 1, 1]
            shape[a] = n3[a]
            bc[(kind, a)] = (torch.tensor(b, **kw).reshape(shape), torch.tensor(b - 1, **kw).reshape(shape))
    inv_eps = torch.ones(n3, **kw)
    i0, j0, k0 = d0
    ni, nj, nk = eps_design.shape
    inv_eps = inv_eps.clone()
    inv_eps[i0:i0 + ni, j0:j0 + nj, k0:k0 + nk] = 1.0 / eps_design.to(inv_eps.device)
    E = [torch.zeros(n3, **kw) for _ in range(3)]
    H = [torch.zeros(n3, **kw) for _ in range(3)]
    # curl terms per component: (axis, field component, sign)
    curl = {0: ((1, 2, 1), (2, 1, -1)), 1: ((2, 0, 1), (0, 2, -1)), 2: ((0, 1, 1), (1, 0, -1))}
    psiH = {(c, a): torch.zeros(n3, **kw) for c in range(3) for (a, _, _) in curl[c]}
    psiE = {(c, a): torch.zeros(n3, **kw) for c in range(3) for (a, _, _) in curl[c]}
    mon = torch.tensor(np.array(monitors), dtype=torch.long, device=device)
    Fr = torch.zeros(len(monitors), **kw)
    Fi = torch.zeros(len(monitors), **kw)
    for n in range(steps):
        Hn = []
        for c in range(3):
            tot = 0
            for (a, src, sg) in curl[c]:
                D = d_bwd(E[src], a, dd[a])
                b, cc = bc[("H", a)]
                psiH[(c, a)] = b * psiH[(c, a)] + cc * D
                tot = tot + sg * (D + psiH[(c, a)])
            Hn.append(H[c] - dt * tot)
        H = Hn
        th = (n + 0.5) * dt
        En = []
        for c in range(3):
            tot = 0
            for (a, src, sg) in curl[c]:
                D = d_fwd(H[src], a, dd[a])
                b, cc = bc[("E", a)]
                psiE[(c, a)] = b * psiE[(c, a)] + cc * D
                tot = tot + sg * (D + psiE[(c, a)])
            for q in range(cfg.nsoft):
                so = cfg.soft[q]
                if so.comp == c:
                    J = so.amp * float(purcell.waveform(so.waveform, so.w, so.kEnv, so.t0, so.tw, th))
                    src = torch.zeros(n3, **kw)
                    src[so.i, so.j, so.k] = J
                    tot = tot - src
            En.append(E[c] + dt * inv_eps * tot)
        E = En
        t = (n + 1) * dt
        ez = E[2][mon[:, 0], mon[:, 1], mon[:, 2]]
        Fr = Fr + ez * math.cos(omega * t) * dt
        Fi = Fi + ez * math.sin(omega * t) * dt
    return (Fr * Fr + Fi * Fi).sum()
