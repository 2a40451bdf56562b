"""The package's Purcell post-processing (dtmaxwell4_amd/purcell.py) against
the oracle's literal purcell.gp:2 expression and waveform, and the spectrum
helper against a direct sum."""
import json
import os

import numpy as np
import pytest

from dtmaxwell4_amd import abi, purcell
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


def test_purcell_factor_is_the_reference_expression():
    """Device spectra F are the conjugates of the reference's: purcell.gp:2
    gets (|F|, -arg F)."""
    rng = np.random.RandomState(1)
    for _ in range(50):
        w = rng.uniform(0.5, 20)
        FE = complex(*rng.normal(size=2)) * 10 ** rng.uniform(-3, 3)
        Fd = complex(*rng.normal(size=2)) * 10 ** rng.uniform(-3, 3)
        ref = orc.purcell(w, abs(FE), -np.angle(FE), abs(Fd), -np.angle(Fd))
        assert purcell.purcell_factor(w, FE, Fd) == pytest.approx(ref, rel=1e-12, abs=1e-12)


def test_reference_spectra_use_the_conjugate_transform():
    """The phase of the reference's fdip.dat falls with w (slope -0.5, the
    pulse centre time with a minus sign): its transform kernel is e^{-iwt}.
    The device-convention spectrum of a pulse centred at t_c rises with slope
    +t_c."""
    data = json.load(open(os.path.join(HERE, "golden", "purcell_rows.json")))
    w = np.array([r["omega"] for r in data["rows"]])
    ph = np.unwrap(np.array([r["phd"] for r in data["rows"]]))
    slope = np.polyfit(w, ph, 1)[0]
    assert slope == pytest.approx(-0.5, abs=0.02)
    t = abi.Tfsf()
    t.waveform, t.w, t.t0, t.tw, t.amp = abi.WAVE_GAUSS_SINE, 12.0, 0.5, 0.2, 1.0
    om = np.linspace(9.0, 15.0, 13)
    F = purcell.dipole_spectrum(t, om, 4000, 1e-3)
    ours = np.polyfit(om, np.unwrap(np.angle(F)), 1)[0]
    assert ours == pytest.approx(0.5, abs=0.02)


@pytest.mark.parametrize("kind", [abi.WAVE_RAISED_COS, abi.WAVE_GAUSS_SINE])
def test_waveform_matches_oracle(kind):
    w, kEnv, t0, tw = 7.0, 0.2, 1.3, 0.4
    t = np.linspace(-0.5, 6.0, 97)
    ours = purcell.waveform(kind, w, kEnv, t0, tw, t)
    ref = np.array([orc.wave(kind, w, kEnv, t0, tw, x) for x in t])
    np.testing.assert_allclose(ours, ref, rtol=1e-12, atol=1e-15)


def test_dipole_spectrum_direct_sum():
    t = abi.Tfsf()
    t.waveform, t.w, t.t0, t.tw, t.amp = abi.WAVE_GAUSS_SINE, 5.0, 1.0, 0.3, 2.0
    dt, steps = 0.01, 400
    om = np.array([4.0, 5.0, 6.5])
    F = purcell.dipole_spectrum(t, om, steps, dt)
    for q, wq in enumerate(om):
        ref = sum(2.0 * orc.wave(t.waveform, t.w, 0.0, t.t0, t.tw, (n + 1) * dt) * np.exp(1j * wq * (n + 1) * dt) * dt
                  for n in range(steps))
        assert F[q] == pytest.approx(ref, rel=1e-12)


def test_reference_dipole_moment_is_the_gaussian_envelope_pulse():
    """The reference's fdip.dat is the spectrum of its EnvelopeTime pulse
    d(t) = exp(-((t - t0)/tw)^2) (signal.hpp:55, no carrier) with the
    TFSFsrc::set values t0 = 25*0.02, tw = 5*0.02 (signal.hpp:99-115): the
    amplitude falls as exp(-w^2 tw^2 / 4) (measured 5e-6 from it over the
    42 rows) and the phase as -w t0.  DTM_WAVE_GAUSS is that pulse: its
    device spectrum has the same shape."""
    data = json.load(open(os.path.join(HERE, "golden", "purcell_rows.json")))
    w = np.array([r["omega"] for r in data["rows"]])
    A = np.array([r["Ad"] for r in data["rows"]])
    ph = np.unwrap(np.array([r["phd"] for r in data["rows"]]))
    tw, t0 = 5 * 0.02, 25 * 0.02
    shape = np.exp(-(w ** 2 - w[0] ** 2) * tw ** 2 / 4)
    assert np.abs(A / A[0] / shape - 1).max() < 2e-5
    assert np.polyfit(w, ph, 1)[0] == pytest.approx(-t0, abs=1e-4)
    # a carrier (the Gaussian-modulated sine) would not fit
    t = abi.Tfsf()
    t.waveform, t.w, t.t0, t.tw, t.amp = abi.WAVE_GAUSS, 2 * np.pi / 0.5, t0, tw, 1.0
    dt = 0.002
    F = purcell.dipole_spectrum(t, w, 4000, dt)
    assert np.abs(np.abs(F) / np.abs(F[0]) / shape - 1).max() < 1e-3
    assert np.polyfit(w, np.unwrap(np.angle(F)), 1)[0] == pytest.approx(t0, abs=1e-3)  # device: e^{+iwt}
