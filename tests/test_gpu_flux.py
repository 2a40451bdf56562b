"""FP64 Poynting-flux sums on the device (dtm_add_flux_box): the kernel's
face sums against a numpy restatement from the fields, the radiated energy
of a vacuum soft dipole against the closed form, and energy in = energy out
for a box the pulse only crosses.  Box energies (dtm_energy_box) against a
numpy sum of the same fields, and reproducible bit for bit."""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs, purcell

pytestmark = pytest.mark.gpu


def solver(cfg):
    from dtmaxwell4_amd.solver import Solver
    assert abi.load_library().dtm_device_count() > 0, "no HIP device visible"
    return Solver(cfg)


def face_flux(E0, E1, H, box, d):
    """Outward flux of the six faces from E^n (E0), E^{n+1} (E1) and
    H^{n+1/2} (H), with the face rule of dtm_flux.hip restated: per face
    normal a (b = a+1, c = a+2), S_a = E_b H_c - E_c H_b, E averaged across
    the face plane and over the two time levels, trapezoid along the even
    tangential coordinate, midpoint along the odd one."""
    lo, hi = box[0::2], box[1::2]
    out = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for side, p in ((0, lo[a]), (1, hi[a])):
            tot = 0.0
            for term, (ce, ch) in enumerate(((b, c), (c, b))):
                nb = hi[b] - lo[b] + (1 if term == 0 else 0)
                nc = hi[c] - lo[c] + (0 if term == 0 else 1)
                wb = np.ones(nb)
                wc = np.ones(nc)
                if term == 0:
                    wb[0] = wb[-1] = 0.5
                else:
                    wc[0] = wc[-1] = 0.5
                sl = [slice(None)] * 3
                sl[b] = slice(lo[b], lo[b] + nb)
                sl[c] = slice(lo[c], lo[c] + nc)
                s1, s0 = list(sl), list(sl)
                s1[a], s0[a] = p, p - 1
                e = 0.25 * (E0[ce][tuple(s0)].astype(np.float64) + E0[ce][tuple(s1)] + E1[ce][tuple(s0)] +
                            E1[ce][tuple(s1)])
                h = H[ch][tuple(s1)].astype(np.float64)
                # the 2-D slices keep (b, c) in increasing axis order
                w = np.outer(wb, wc) if b < c else np.outer(wb, wc).T
                v = float((e * h * w).sum())
                tot += v if term == 0 else -v
            out.append(tot * d[b] * d[c] * (1 if side else -1))
    return np.array(out)


def dipole(n=64, npml=10):
    w = configs.vacuum_dipole(n=n, npml=npml, steps=0)
    return w


def test_flux_sums_match_numpy_restatement():
    w = dipole(n=48, npml=8)
    s = solver(w.cfg)
    box = (18, 31, 17, 30, 19, 29)
    s.add_flux_box(*box)
    s.step(300)
    E0 = [s.field(c) for c in range(3)]
    s.step(1)
    E1 = [s.field(c) for c in range(3)]
    H = [s.field(c) for c in range(3, 6)]
    got = s.flux()[-1, 0]
    ref = face_flux(E0, E1, H, box, (w.cfg.dx, w.cfg.dy, w.cfg.dz))
    print("kernel", got, "numpy", ref)
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
    s.close()


def test_radiated_energy_of_a_vacuum_dipole_matches_closed_form():
    """A soft current J(t) in one cell radiates P = (dV dJ/dt)^2 / (6 pi)
    (Larmor, c = mu = eps = 1): the outward flux through a box 12 cells
    around it, integrated over the run on the device in FP64, gives the
    radiated energy."""
    w = dipole()
    cfg = w.cfg
    s = solver(cfg)
    c = cfg.nx // 2
    s.add_flux_box(c - 12, c + 13, c - 12, c + 13, c - 12, c + 13)
    steps = 1700
    s.step(steps)
    W = float(s.flux_total()[0].sum())
    so = cfg.soft[0]
    t = np.linspace(0, steps * cfg.dt, 40001)
    J = so.amp * np.array([purcell.waveform(so.waveform, so.w, so.kEnv, so.t0, so.tw, x) for x in t])
    dJ = np.gradient(J, t)
    dV = cfg.dx * cfg.dy * cfg.dz
    W_th = float(np.trapezoid((dV * dJ) ** 2, t) / (6 * np.pi))
    # the pulse has left the box: the flux series ends near zero
    fl = s.flux()[:, 0].sum(1)
    print("radiated", W, "closed form", W_th, "ratio", W / W_th, "last/peak flux", abs(fl[-1]) / np.abs(fl).max())
    assert abs(fl[-1]) < 1e-3 * np.abs(fl).max()
    assert abs(W / W_th - 1) < 0.02
    s.close()


def test_box_crossed_by_a_pulse_gets_back_what_enters():
    """A box beside the dipole holds no source: once the pulse has crossed
    it, the net outward energy is zero to the second-order face rule."""
    w = dipole()
    cfg = w.cfg
    s = solver(cfg)
    c = cfg.nx // 2
    s.add_flux_box(c + 4, c + 16, c - 6, c + 6, c - 6, c + 6)
    s.step(1700)
    tot = s.flux_total()[0]
    inflow = -tot[0]  # through the face that looks at the dipole
    print("per face", tot, "net / inflow", tot.sum() / inflow)
    assert inflow > 0
    assert abs(tot.sum()) < 5e-3 * inflow
    s.close()


def test_box_energy_matches_numpy_and_is_reproducible():
    w = configs.plane_wave_sphere(nx=64, ny=48, nz=48, npml=8, inset=4, radius=8, steps=200, periods_w=1.0,
                                  periods_t0=2.0)
    s = configs.apply(w, solver(w.cfg))
    s.step(200)
    f = [x.astype(np.float64) for x in s.fields()]
    eps = [np.where(m == 1, w.cfg.mat[1].eps_inf, 1.0) for m in w.mats]
    box = (10, 50, 7, 41, 5, 44)
    sl = (slice(10, 50), slice(7, 41), slice(5, 44))
    dv = w.cfg.dx * w.cfg.dy * w.cfg.dz
    ref = 0.5 * dv * sum(float((eps[c][sl] * f[c][sl] ** 2).sum()) + float((f[3 + c][sl] ** 2).sum())
                         for c in range(3))
    e1 = s.energy(box)
    e2 = s.energy(box)
    print("box energy", e1, "numpy", ref)
    assert e1 == e2
    assert abs(e1 / ref - 1) < 1e-6
    full = 0.5 * dv * sum(float((eps[c] * f[c] ** 2).sum()) + float((f[3 + c] ** 2).sum()) for c in range(3))
    assert abs(s.energy() / full - 1) < 1e-6
    s.close()


def test_flux_box_rejections():
    w = dipole(n=32, npml=6)
    s = solver(w.cfg)
    with pytest.raises(abi.DTMError):
        s.add_flux_box(0, 10, 5, 10, 5, 10)     # a face on plane 0
    with pytest.raises(abi.DTMError):
        s.add_flux_box(5, 32, 5, 10, 5, 10)     # a face on plane n
    with pytest.raises(abi.DTMError):
        s.add_flux_box(5, 5, 5, 10, 5, 10)      # empty
    s.step(1)
    with pytest.raises(abi.DTMError):
        s.add_flux_box(5, 10, 5, 10, 5, 10)     # after the first step
    s.close()
