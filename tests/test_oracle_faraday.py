"""Pins the magnetised-plasma current (the build's gyrotropic model; the
reference's Gyration() is disabled, genAlg.py:303-304) against Faraday
rotation: a slab with dJ/dt + nu J = wp^2 E + wc x J, bias along z, has
circular eigenmodes E_+- = Ex +- i Ey with
    eps_+- = eps_inf - wp^2 / (w (w +- wc + i nu))   (e^{-iwt} convention),
each transmitted through the slab with the Fabry-Perot coefficient
    t = 4 n e^{i(n-1)kL} / ((1+n)^2 - (1-n)^2 e^{2inkL}).
The x-polarised incident pulse leaves the slab rotated: the measured ratio
t_+/t_- (transmitted circular components over the incident, at one
frequency) must match the closed form.  A wrong sign of the cyclotron term
swaps the modes and inverts the ratio's phase; a wrong wp or wc moves it."""
import numpy as np

from dtmaxwell4_amd import abi, configs
from oracle.oracle import Oracle


def transmitted(w, with_slab):
    o = Oracle(w.cfg)
    if with_slab:
        o.set_materials(*w.mats)
    for p in w.probes:
        o.add_probe(*p)
    o.set_dft(w.omegas)
    o.step(w.steps)
    F = o.dft()[:, 0]
    o.close()
    return F[0], F[1]  # Ex, Ey at the probe, e^{+iwt} transform = e^{-iwt} amplitude


def slab_t(n, k, L):
    return 4 * n * np.exp(1j * (n - 1) * k * L) / ((1 + n) ** 2 - (1 - n) ** 2 * np.exp(2j * n * k * L))


def test_faraday_rotation_of_magnetised_plasma_slab():
    w = configs.gyro_slab(nx=4, ny=4, nz=256, npml=12, thick=64, steps=6000, cells_per_lambda=40)
    m = w.cfg.mat[1]
    om = float(w.omegas[0])
    Fx, Fy = transmitted(w, True)
    Fx0, Fy0 = transmitted(w, False)
    assert abs(Fy0) < 1e-6 * abs(Fx0)                 # incident is x-polarised
    tp = (Fx + 1j * Fy) / Fx0
    tm = (Fx - 1j * Fy) / Fx0
    wp, nu, wc = m.gyro_wp, m.gyro_nu, m.gyro_wc[2]
    eps = {s: m.eps_inf - wp ** 2 / (om * (om + s * wc + 1j * nu)) for s in (1, -1)}
    n = {s: np.sqrt(eps[s] + 0j) for s in (1, -1)}
    L = w.notes["thick"] * w.notes["dx"]
    ratio_th = slab_t(n[1], om, L) / slab_t(n[-1], om, L)
    ratio = tp / tm
    print("t+/t- measured", ratio, "closed form", ratio_th)
    assert abs(n[1] - n[-1]) > 0.5                    # strongly gyrotropic: a real test
    assert abs(ratio - ratio_th) < 0.05 * abs(ratio_th)
    # the rotation is not trivial
    assert abs(np.angle(ratio_th)) > 0.3
