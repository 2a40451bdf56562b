"""The gyrotropic workload at its full size (256x256x512, periodic in x and
y, 64-cell magnetised-plasma slab with eps_inf 4, wp 2e15 rad/s, nu 1e13
rad/s, wc 0.2 wp along z, normally incident x-polarised TF/SF pulse, 8000
steps) on the GPU, against the closed-form Faraday rotation of the slab.

The circular eigenmodes E_+- = Ex +- i Ey see
    eps_+- = eps_inf - wp^2 / (w (w +- wc + i nu))   (e^{-iwt} convention)
and cross the slab with the Fabry-Perot coefficient
    t = 4 n e^{i(n-1)kL} / ((1+n)^2 - (1-n)^2 e^{2inkL}),
so the transmitted circular components over the incident one (DFT at the
pulse centre frequency, probe behind the slab, same probe without the slab)
give t_+/t_-.  The grid is uniform across x and y, so the full-size run
sees the same one-dimensional physics as tests/test_oracle_faraday.py."""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs

pytestmark = pytest.mark.gpu


def transmitted(w, with_slab):
    from dtmaxwell4_amd.solver import Solver
    assert abi.load_library().dtm_device_count() > 0, "no HIP device visible"
    s = Solver(w.cfg)
    if with_slab:
        s.set_materials(*w.mats)
    for p in w.probes:
        s.add_probe(*p)
    s.set_dft(w.omegas)
    s.step(w.steps)
    F = s.dft()[:, 0]
    s.close()
    return F[0], F[1]


def slab_t(n, k, L):
    return 4 * n * np.exp(1j * (n - 1) * k * L) / ((1 + n) ** 2 - (1 - n) ** 2 * np.exp(2j * n * k * L))


def test_full_size_faraday_rotation_matches_closed_form():
    w = configs.gyro_slab()
    m = w.cfg.mat[1]
    om = float(w.omegas[0])
    Fx, Fy = transmitted(w, True)
    Fx0, Fy0 = transmitted(w, False)
    assert abs(Fy0) < 1e-5 * abs(Fx0)  # incident is x-polarised
    tp = (Fx + 1j * Fy) / Fx0
    tm = (Fx - 1j * Fy) / Fx0
    wp, nu, wc = m.gyro_wp, m.gyro_nu, m.gyro_wc[2]
    eps = {s: m.eps_inf - wp ** 2 / (om * (om + s * wc + 1j * nu)) for s in (1, -1)}
    n = {s: np.sqrt(eps[s] + 0j) for s in (1, -1)}
    L = w.notes["thick"] * w.notes["dx"]
    ratio_th = slab_t(n[1], om, L) / slab_t(n[-1], om, L)
    ratio = tp / tm
    print("t+/t- measured", ratio, "closed form", ratio_th)
    assert abs(n[1] - n[-1]) > 0.5  # strongly gyrotropic
    assert abs(np.angle(ratio_th)) > 0.3  # a rotation well away from zero
    assert abs(ratio - ratio_th) < 0.05 * abs(ratio_th)
