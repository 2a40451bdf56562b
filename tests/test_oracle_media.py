"""Pins the oracle's absorbing layer and media against continuous physics:

* CPML sigma grading (DTmxw.cu:220-249): the round-trip attenuation of the
  continuous profile equals the target R exactly, so the discretely
  integrated sigma must give R (pins sigma_max's factor (m+1)/(2L) and ln R);
* CPML absorbs: a pulse in a short box agrees with a long box (no
  reflection) to the discretisation level; so does the reference's sigma
  grading (DTM_PML_PROFILE_REFERENCE) in the CPML and the split-field layer,
  whose flat outer and zero inner margins are pinned on the profile itself;
* the Lorentz/Drude ADE recursion (dispersion.cpp:41-99) has a discrete
  transfer function that converges at second order to the continuous
  chi(w) = (a0 - i w a1)/(b0 - i w b1 - w^2) and eps(w) = eps_inf + chi + i sigma/w;
* the magnetised-plasma current update converges at second order to
  dJ/dt + nu J = wp^2 E + wc x J, and its rotation part is lossless;
* CoffStruct::set_eps Cayley identities (defs.h:96-117).
"""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs
from oracle import oracle as orc
from oracle.oracle import Oracle


def test_pml_sigma_integrates_to_target_reflection():
    for m, R, npml, d in [(4.0, 1e-4, 12, 0.05), (3.0, 1e-6, 20, 0.02), (4.0, 1e-4, 10, 0.005)]:
        n = 100
        h = np.arange(0, 2 * npml + 1)
        s = np.array([orc.pml_sigma(int(x), n, npml, npml, d, m, R, 1.0) for x in h])
        integral = np.trapezoid(s, h * d / 2)       # over the slab, half-cell samples
        assert np.exp(-2 * integral) == pytest.approx(R, rel=0.05)
        # graded: maximal at the outer wall, zero at the inner face
        assert s[0] == s.max() and s[-1] == 0.0
        # high side mirrors the low side
        sh = np.array([orc.pml_sigma(int(2 * n - x), n, npml, npml, d, m, R, 1.0) for x in h])
        np.testing.assert_allclose(sh, s)


@pytest.mark.parametrize("p", [4, 8, 15, 22])
def test_reference_sigma_profile(p):
    """DTM_PML_PROFILE_REFERENCE is PMLgamma_func (DTmxw.cu:219-228) as
    setPMLcoeffs indexes it along x (DTmxw.cu:251-253): sigma_max over the
    outer 3 half-cells (the clamp of i - 3 < 0), zero over the inner 3 (the
    early return at i >= N - 3), and between them the m-th power grading
    whose round trip gives exactly R; both sides of an axis alike."""
    n, d, m, R, V = 100, 0.01, 4.0, 1e-4, 1.0
    u = np.arange(0, 2 * p + 1)
    s = np.array([orc.pml_sigma(int(x), n, p, p, d, m, R, V, 1) for x in u])
    N = 2 * p - 6
    assert np.all(s[:4] == s[0]) and s[0] == s.max() and s[0] > 0
    assert np.all(s[2 * p - 3:] == 0.0) and s[2 * p - 4] > 0
    assert np.all(np.diff(s) <= 0)
    if N >= 10:  # the graded part's round trip, -2 int sigma / V = ln R (trapezoid on half-cells)
        integral = np.trapezoid(s[3:3 + N + 1], dx=d / 2)
        assert 2 * integral / V == pytest.approx(np.log(1 / R), rel=0.03)
    hi = np.array([orc.pml_sigma(int(2 * n - x), n, p, p, d, m, R, V, 1) for x in u])
    np.testing.assert_array_equal(hi, s)
    # the default grading is untouched by the option
    assert orc.pml_sigma(0, n, p, p, d, m, R, V, 0) == orc.pml_sigma(0, n, p, p, d, m, R, V)


def pulse_run(nx, steps, npml, kind=abi.PML_CPML, profile=abi.PML_PROFILE_POLY):
    d = 0.05
    cfg = abi.default_config(nx, 4, 4, d, d, d, configs.dt_for(d, d, d, 0.5))
    cfg.bc[1] = cfg.bc[2] = abi.BC_PERIODIC
    cfg.npml[0] = cfg.npml[1] = npml
    cfg.pml_kind, cfg.pml_profile = kind, profile
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EZ, nx // 2, 0, 0
    so.waveform, so.amp, so.w, so.tw, so.t0 = abi.WAVE_GAUSS_SINE, 1.0, 2 * np.pi, 0.3, 1.0
    o = Oracle(cfg)
    probe_i = nx // 2 + 30
    o.add_probe(abi.EZ, probe_i, 0, 0)
    o.step(steps)
    return o.probes()[:, 0]


def test_cpml_absorbs_outgoing_pulse():
    steps = 900
    short = pulse_run(140, steps, 16)   # boundary 40 cells beyond the probe
    long_ = pulse_run(600, steps, 16)   # boundary far away: no echo within the window
    peak = np.abs(long_).max()
    assert peak > 0
    refl = np.abs(short - long_).max() / peak
    assert refl < 5e-3


@pytest.mark.parametrize("kind", [abi.PML_CPML, abi.PML_SPLIT])
def test_reference_profile_absorbs_outgoing_pulse(kind):
    """The reference's sigma grading absorbs in either layer kind: its 3-cell
    flat and zero margins cost little against the default grading."""
    steps = 900
    long_ = pulse_run(600, steps, 16)
    short = pulse_run(140, steps, 16, kind, abi.PML_PROFILE_REFERENCE)
    refl = np.abs(short - long_).max() / np.abs(long_).max()
    print("reflection", refl)
    assert refl < 1e-2


def ade_transfer(mat, dt, w):
    """Discrete response of the oracle's coefficients to E^n = e^{-i w n dt}:
    returns (J/E, eps_eff) from the J and E recursions (dispersion.cpp:85-99)."""
    kE, kEJ, kJ = orc.disp_coeffs(mat, dt)
    z = np.exp(-1j * w * dt)
    Jsum = 0.0
    Jterm = 0.0
    for p in range(mat.npoles):
        a, xi, zp, zm, z0 = kJ[p]
        # J^{n+1} = a J^n + xi J^{n-1} + zp E^{n+1} + zm E^{n-1} + z0 E^n  (coefficients already /dt)
        Jp = (zp * z + zm / z + z0) / (z - a - xi / z)
        Jsum += Jp
        Jterm += kEJ[p][0] * Jp + kEJ[p][1] * Jp / z
    # E^{n+1} = kE0 E^{n-1} + kE1 E^n + kE2 (dt curl^{n+1/2}) - sum(kEJ0 J^n + kEJ1 J^{n-1})
    # => curl^{n+1/2} amplitude C with curl = C e^{-i w (n+1/2) dt}:
    C = (z - kE[0] / z - kE[1] + Jterm) / (kE[2] * dt * np.sqrt(z))
    wt = 2 / dt * np.sin(w * dt / 2)
    eps_eff = C / (-1j * wt)                 # curl H = dD/dt -> -i w~ eps E
    return Jsum, eps_eff


def continuous_eps(mat, w):
    chi = 0
    for p in range(mat.npoles):
        L = mat.pole[p]
        chi += (L.a0 - 1j * w * L.a1) / (L.b0 - 1j * w * L.b1 - L.b2 * w * w)
    return mat.eps_inf + chi + 1j * mat.sigma / w


@pytest.mark.parametrize("kind", ["drude", "lorentz", "both_lossy"])
def test_ade_discrete_permittivity_converges_second_order(kind):
    m = abi.Material()
    m.eps_inf = 3.7
    if kind == "drude":
        m.npoles = 1
        m.pole[0].setDrude(1.0, 45.7, 0.045, 0.0)
    elif kind == "lorentz":
        m.npoles = 1
        m.pole[0].set(0.5, 25.0, 1.25, 0.0)
    else:
        m.npoles = 2
        m.pole[0].setDrude(1.0, 45.7, 0.045, 0.3)
        m.pole[1].set(0.5, 25.0, 1.25, 0.0)
        m.sigma = 0.7
    w = 10.5
    errs = []
    for dt in (0.004, 0.002, 0.001):
        _, eps = ade_transfer(m, dt, w)
        errs.append(abs(eps - continuous_eps(m, w)) / abs(continuous_eps(m, w)))
    assert errs[-1] < 1e-3
    assert errs[0] / errs[1] == pytest.approx(4, rel=0.25)
    assert errs[1] / errs[2] == pytest.approx(4, rel=0.25)


def test_gyro_current_update_converges_and_rotation_is_lossless():
    m = abi.Material()
    m.gyro, m.gyro_wp, m.gyro_nu = 1, 6.7, 0.03
    m.gyro_wc[:] = [0.3, -0.5, 1.34]
    w = 4.0
    errs = []
    for dt in (0.02, 0.01, 0.005):
        A, B = orc.gyro_coeffs(m, dt)
        # J^{n+1/2} = A J^{n-1/2} + B E^n with E^n = E0 e^{-iwn dt}, J = J0 e^{-iw(n+1/2)dt}
        z = np.exp(-1j * w * dt)
        E0 = np.array([1.0, 0.2, -0.4])
        J = np.linalg.solve(np.sqrt(z) * np.eye(3) - A / np.sqrt(z), B @ E0)
        # continuous: (-iw + nu) J - wc x J = wp^2 E
        W = np.array([[0, -m.gyro_wc[2], m.gyro_wc[1]], [m.gyro_wc[2], 0, -m.gyro_wc[0]],
                      [-m.gyro_wc[1], m.gyro_wc[0], 0]])
        Jc = np.linalg.solve((-1j * w + m.gyro_nu) * np.eye(3) - W, m.gyro_wp ** 2 * E0)
        errs.append(np.linalg.norm(J - Jc) / np.linalg.norm(Jc))
    assert errs[0] / errs[1] == pytest.approx(4, rel=0.3)
    assert errs[1] / errs[2] == pytest.approx(4, rel=0.3)
    m.gyro_nu = 0.0
    A, _ = orc.gyro_coeffs(m, 0.01)
    np.testing.assert_allclose(np.abs(np.linalg.eigvals(A)), 1.0, atol=1e-12)


def test_set_eps_cayley_identities():
    """deps*(2 eps + G) = 2I and gyration = (eps+G/2)^-1 (eps-G/2), G the
    antisymmetric matrix of g*w*dt, hold exactly when gx = 0."""
    eps, w, dt = 5.0, 3.0, 0.002
    for g in [(0.0, -0.05, 0.1), (0.0, 0.0, 0.3), (0.0, 0.2, 0.0)]:
        D, Gy = orc.set_eps(eps, *g, w, dt)
        gx, gy, gz = [v * w * dt for v in g]
        G = np.array([[0, -gz, gy], [gz, 0, -gx], [-gy, gx, 0]])
        np.testing.assert_allclose(D @ (2 * eps * np.eye(3) + G), 2 * np.eye(3), atol=1e-12)
        np.testing.assert_allclose(Gy, np.linalg.solve(eps * np.eye(3) + G / 2, eps * np.eye(3) - G / 2),
                                   atol=1e-12)


def test_set_eps_reference_xz_term():
    """For gx != 0 the reference's coffAnisoXZ = (gx*gy - 2 epsY gy)/D
    (defs.h:102) differs from the inverse's (gx*gz - 2 epsY gy)/D by
    gx (gy - gz)/D; every other entry is the exact inverse.  The restatement
    keeps the reference's expression (gyration is disabled in the reference's
    kernels, genAlg.py:303-304, so no result depends on it)."""
    eps, w, dt = 5.0, 3.0, 0.002
    g = (0.02, -0.05, 0.1)
    D, _ = orc.set_eps(eps, *g, w, dt)
    gx, gy, gz = [v * w * dt for v in g]
    G = np.array([[0, -gz, gy], [gz, 0, -gx], [-gy, gx, 0]])
    exact = 2 * np.linalg.inv(2 * eps * np.eye(3) + G)
    Dd = 4 * eps ** 3 + eps * (gx * gx + gy * gy + gz * gz)
    diff = D - exact
    assert diff[0, 2] == pytest.approx(gx * (gy - gz) / Dd, rel=1e-9)
    diff[0, 2] = 0
    np.testing.assert_allclose(diff, 0, atol=1e-15)
