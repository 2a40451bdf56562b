"""Purcell factor of a dipole over a mirror, end to end on the FP64 oracle:
the reference's DipoleTFSF (scattered field inside its box, signal.hpp:51-97)
for an Ez dipole kd cells above the zero boundary at the low-z face, CPML on
the other faces, the device DFT convention at the dipole, and the package's
purcell_factor (purcell.gp:2).  Fields beyond the grid are zero, so the
tangential E vanishes at z = -dx/2 (Ex, Ey of cell k = -1): a perfect mirror
a height h = (kd + 1/2) dx below the dipole, for which
P = 1 + 3 (sin x / x^3 - cos x / x^2), x = 2 k h.  A transform-sign mistake
moves P - 1 to the other side of zero (0.79 instead of 1.22 at w0, kd = 5)."""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs, purcell
from oracle.oracle import Oracle


@pytest.mark.parametrize("kd", [5, 8])
def test_dipole_over_mirror_matches_image_closed_form(kd):
    n, lam = 56, 20
    dx = 1.0 / lam
    dt = configs.dt_for(dx, dx, dx)
    cfg = abi.default_config(n, n, n, dx, dx, dx, dt)
    cfg.npml[:] = [10, 10, 10, 10, 0, 10]
    ci, bh = n // 2, 2
    pos = np.array([(ci + 0.5) * dx, (ci + 0.5) * dx, kd * dx])  # Ez node of cell (ci, ci, kd)
    t = cfg.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_DIPOLE, abi.WAVE_GAUSS_SINE, 1
    h = bh * dx
    t.box[:] = [pos[0] - h, pos[0] + h, pos[1] - h, pos[1] + h, pos[2] - h, pos[2] + h]
    w0 = 2 * np.pi
    t.w, t.cSrc, t.Imp, t.amp = w0, 1.0, 1.0, h ** 3
    t.tw, t.t0 = 0.5, 2.0
    t.tStop = 1e30
    om = w0 * np.linspace(0.7, 1.3, 7)
    steps = 1200
    o = Oracle(cfg)
    o.add_probe(abi.EZ, ci, ci, kd)
    o.set_dft(om)
    o.step(steps)
    FE = o.dft()[0]
    o.close()
    P = purcell.purcell_factor(om, FE, purcell.dipole_spectrum(t, om, steps, dt))
    x = 2 * om * (kd + 0.5) * dx
    closed = 1 + 3 * (np.sin(x) / x ** 3 - np.cos(x) / x ** 2)
    print("P", P, "closed form", closed)
    assert np.max(np.abs(P - closed)) < 0.03
