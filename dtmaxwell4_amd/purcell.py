"""Purcell-factor post-processing of a dipole TF/SF run.

The reference computes spectra with an external Fourier script
(calc_purcell.sh:2) and the Purcell factor with gnuplot (purcell.gp:2):

    P(w) = (A_E cos(ph_E) A_d sin(ph_d) - A_E sin(ph_E) A_d cos(ph_d)) * w/2 / (A_d^2 w^4 / 3) + 1

with (A_E, ph_E) the spectrum of the scattered field at the dipole (the box
interior is the scattered-field region, signal.hpp:153-160) and (A_d, ph_d)
that of the dipole moment.  The spectra it reads come from a transform with
the kernel e^{-i w t}: the phase of the reference's own fdip.dat falls
linearly with w (slope -0.5, minus the centre time of its pulse), where an
e^{+i w t} transform of a pulse at positive times rises.  Here both spectra
use the device DFT convention F(w) = sum_n f(t_n) e^{+i w t_n} dt (E sampled
at t_n = (n+1) dt), whose values are the complex conjugates of the
reference's: (A, ph) = (|F|, -arg F).
"""
import numpy as np

from . import abi


def waveform(kind, w, kEnv, t0, tw, t):
    """Dipole-moment time function f0 (signal.hpp:58-77 raised cosine, or the
    Gaussian-modulated sine), at phase w t (the dipole's own position, R = 0)."""
    t = np.asarray(t, dtype=np.float64)
    p = w * t
    if kind == abi.WAVE_RAISED_COS:
        kp = p * kEnv
        on = (kp >= 0) & (kp <= 2 * np.pi)
        return np.where(on, np.sin(p) * 0.5 * (1 - np.cos(kp)), 0.0)
    u = (p / w - t0) / tw
    if kind == abi.WAVE_GAUSS:  # EnvelopeTime of signal.hpp:55 (no carrier)
        return np.exp(-u * u)
    return np.exp(-u * u) * np.sin(p)


def dipole_spectrum(tfsf, omegas, steps, dt):
    """F_d(w) = sum_{n<steps} d((n+1) dt) e^{i w (n+1) dt} dt, d = amp * f0."""
    t = (np.arange(steps) + 1.0) * dt
    d = tfsf.amp * waveform(tfsf.waveform, tfsf.w, tfsf.kEnv, tfsf.t0, tfsf.tw, t)
    ph = np.exp(1j * np.outer(np.asarray(omegas, dtype=np.float64), t))
    return (ph * d[None, :]).sum(1) * dt


def purcell_factor(omegas, FE, Fd):
    """purcell.gp:2 on device-convention spectra: with the reference's
    spectra conj(F), its A_E A_d sin(ph_d - ph_E) becomes Im(F_E conj(F_d)),
    so P = 1 + (3/2) Im(F_E conj(F_d)) / (|F_d|^2 w^3)."""
    w = np.asarray(omegas, dtype=np.float64)
    FE = np.asarray(FE)
    Fd = np.asarray(Fd)
    return 1.0 + 1.5 * np.imag(FE * np.conj(Fd)) / (np.abs(Fd) ** 2 * w ** 3)
