"""FP32 zone of FP16 storage (dtm_set_fp32_zone): a box of cells whose six
components every kernel reads and writes in FP32.

Checks, each sharp against a missed zone read or write (an FP16 rounding is
5e-4 of the value, the tolerances below are 1e-6):
- a zone over the whole grid reproduces FP32 storage with every kernel
  involved: CPML or the split-field layer, the dipole TF/SF box, Drude +
  Lorentz ADE cells, a soft source, probes and their FP64 DFT;
- a zone inside the grid, with random initial fields confined to it for the
  steps run (the stencil spreads a field 3 cells a step), reproduces FP32
  storage in every cell, at zone edges that cut the sweep tiles in y and z,
  with zero and with periodic boundaries;
- the Purcell factor of the dispersive dipole run with the zone around the
  dipole box tracks FP32 storage to 1e-3, where FP16 storage alone is off by
  tens of percent (rounding of the large total field just outside the box);
- save / load carries the zone; misuse is refused."""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs, purcell

pytestmark = pytest.mark.gpu


def _solver(cfg):
    from dtmaxwell4_amd.solver import Solver
    return Solver(cfg)


def _fields(s):
    return [s.field(c) for c in range(6)]


def _maxrel(a, b):
    den = max(float(np.max(np.abs(y))) for y in b)
    return max(float(np.max(np.abs(x.astype(np.float64) - y))) for x, y in zip(a, b)) / den


def purcell_case(storage, steps=1500):
    w = configs.drude_lorentz_purcell(n=48, npml=8, radius=8, gap=3, nfreq=16, steps=steps,
                                      cells_per_lambda=20, storage=storage)
    cfg = w.cfg
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EX, 14, 30, 33
    so.waveform, so.amp, so.w = abi.WAVE_GAUSS_SINE, 1e-3, w.notes["w0"]
    so.tw, so.t0 = cfg.tfsf.tw, cfg.tfsf.t0
    return w


@pytest.mark.parametrize("pml", [abi.PML_CPML, abi.PML_SPLIT])
def test_whole_grid_zone_equals_fp32_storage(pml):
    w32 = purcell_case(abi.STORE_FP32, steps=400)
    w32.cfg.pml_kind = pml
    a = configs.apply(w32, _solver(w32.cfg))
    a.step(w32.steps)
    w16 = purcell_case(abi.STORE_FP16, steps=400)
    w16.cfg.pml_kind = pml
    b = _solver(w16.cfg)
    n = w16.cfg.nx
    b.set_fp32_zone(0, n, 0, n, 0, n)
    configs.apply(w16, b)
    b.step(w16.steps)
    e_f = _maxrel(_fields(b), _fields(a))
    pa, pb = a.probes(), b.probes()
    e_p = float(np.max(np.abs(pb - pa)) / np.max(np.abs(pa)))
    da, db = a.dft(), b.dft()
    e_d = float(np.max(np.abs(db - da)) / np.max(np.abs(da)))
    print("fields", e_f, "probes", e_p, "dft", e_d)
    assert e_f < 1e-6 and e_p < 1e-6 and e_d < 1e-6
    a.close()
    b.close()


@pytest.mark.parametrize("periodic", [False, True])
def test_confined_field_in_a_partial_zone_equals_fp32_storage(periodic):
    nx, ny, nz, d = 40, 56, 192, 0.05
    steps = 3
    if periodic:  # the zone spans y and z: the wrapped halos come from it
        zone = (4, 36, 0, ny, 0, nz)
        blob = (16, 24, 0, ny, 0, nz)
    else:          # zone edges inside row tiles (y) and z tiles (64 cells)
        zone = (4, 36, 5, 51, 30, 150)
        blob = (16, 24, 17, 39, 42, 138)
    rng = np.random.default_rng(11)
    init = []
    for c in range(6):
        f = np.zeros((nx, ny, nz), np.float32)
        i0, i1, j0, j1, k0, k1 = blob
        f[i0:i1, j0:j1, k0:k1] = rng.standard_normal((i1 - i0, j1 - j0, k1 - k0)).astype(np.float32)
        init.append(f)
    out = {}
    for store in (abi.STORE_FP32, abi.STORE_FP16):
        cfg = abi.default_config(nx, ny, nz, d, d, d, configs.dt_for(d, d, d))
        cfg.storage = store
        if periodic:
            cfg.bc[:] = [abi.BC_PERIODIC] * 3
        if store == abi.STORE_FP16:
            cfg.scale_e_log2[:] = [8] * 3
            cfg.scale_h_log2[:] = [8] * 3
        s = _solver(cfg)
        if store == abi.STORE_FP16:
            s.set_fp32_zone(*zone)
        for c in range(6):
            s.set_field(c, init[c])
        s.step(steps)
        out[store] = _fields(s)
        s.close()
    a, b = out[abi.STORE_FP32], out[abi.STORE_FP16]
    i0, i1, j0, j1, k0, k1 = zone
    for c in range(6):  # nothing reached the cells outside the zone
        m = np.ones((nx, ny, nz), bool)
        m[i0:i1, j0:j1, k0:k1] = False
        assert not a[c][m].any() and not b[c][m].any()
    e = _maxrel(b, a)
    print("periodic" if periodic else "zero", "max rel", e)
    assert e < 1e-6


def _purcell(s, w):
    FE = s.dft()[0]
    Fd = purcell.dipole_spectrum(w.cfg.tfsf, w.omegas, w.steps, w.cfg.dt)
    return purcell.purcell_factor(w.omegas, FE, Fd)


@pytest.mark.parametrize("margin", [8, 7])  # 7: the zone box starts at an odd z (pairs realign it)
def test_zone_around_the_dipole_brings_fp16_purcell_to_fp32(margin):
    curves = {}
    for name, store, zone in (("fp32", abi.STORE_FP32, False), ("fp16", abi.STORE_FP16, False),
                              ("fp16_zone", abi.STORE_FP16, True)):
        w = configs.drude_lorentz_purcell(n=48, npml=8, radius=8, gap=3, nfreq=16, steps=1500,
                                          cells_per_lambda=20, storage=store)
        s = _solver(w.cfg)
        if zone:
            s.set_fp32_zone(*configs.dipole_zone(w, margin))
        configs.apply(w, s)
        s.step(w.steps)
        curves[name] = _purcell(s, w)
        s.close()
    ref = curves["fp32"]
    e16 = float(np.max(np.abs(curves["fp16"] - ref)) / np.max(np.abs(ref)))
    ez = float(np.max(np.abs(curves["fp16_zone"] - ref)) / np.max(np.abs(ref)))
    print("fp16", e16, "fp16 + zone", ez)
    assert ez < 1e-3
    assert ez < 0.1 * e16


def test_state_file_carries_the_zone(tmp_path):
    w = purcell_case(abi.STORE_FP16, steps=60)
    zone = configs.dipole_zone(w)

    def make():
        s = _solver(w.cfg)
        s.set_fp32_zone(*zone)
        return configs.apply(w, s)

    a = make()
    a.step(30)
    p = str(tmp_path / "zone.state")
    a.save_state(p)
    a.step(30)
    b = make()
    b.load_state(p)
    b.step(30)
    for x, y in zip(_fields(a), _fields(b)):
        assert np.array_equal(x, y)
    assert np.array_equal(a.dft(), b.dft())
    a.close()
    b.close()


def test_zone_misuse_is_refused():
    w = purcell_case(abi.STORE_FP32, steps=1)
    s = _solver(w.cfg)
    with pytest.raises(abi.DTMError):   # FP32 storage
        s.set_fp32_zone(0, 8, 0, 8, 0, 8)
    s.close()
    w = purcell_case(abi.STORE_FP16, steps=1)
    n = w.cfg.nx
    s = _solver(w.cfg)
    with pytest.raises(abi.DTMError):   # outside the grid
        s.set_fp32_zone(0, n + 1, 0, 8, 0, 8)
    with pytest.raises(abi.DTMError):   # empty
        s.set_fp32_zone(4, 4, 0, 8, 0, 8)
    s.set_fp32_zone(0, 8, 0, 8, 0, 8)
    with pytest.raises(abi.DTMError):   # twice
        s.set_fp32_zone(0, 8, 0, 8, 0, 8)
    s.close()
    s = _solver(w.cfg)
    s.step(1)
    with pytest.raises(abi.DTMError):   # after the first step
        s.set_fp32_zone(0, 8, 0, 8, 0, 8)
    s.close()
    w.cfg.sweep = abi.SWEEP_FUSED
    s = _solver(w.cfg)
    with pytest.raises(abi.DTMError):   # fused step
        s.set_fp32_zone(0, 8, 0, 8, 0, 8)
    s.close()
