"""Pins the oracle's split-field absorbing layer (genAlg.py:117-148,
coefficients DTmxw.cu:250-265) with checks a wrong sign, a swapped part or
a wrong coefficient would fail:

* with sigma = 0 (R = 1: ln(1/R) = 0) the split update is exactly the plain
  update (the parts sum back to the field);
* the layer absorbs: a pulse in a short box agrees with a long box to the
  discretisation level, like the CPML;
* split-field and CPML agree with each other (both approximate the same
  continuous layer) far better than either agrees with no layer;
* a single-cell split update matches its closed form k1 p + k2 dt dif.
"""
import copy

import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs
from oracle.oracle import Oracle


def pulse(nx, steps, npml, kind, R=1e-4):
    d = 0.05
    cfg = abi.default_config(nx, 4, 4, d, d, d, configs.dt_for(d, d, d, 0.5))
    cfg.bc[1] = cfg.bc[2] = abi.BC_PERIODIC
    cfg.npml[0] = cfg.npml[1] = npml
    cfg.pml_kind = kind
    cfg.pml_R = R
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EZ, nx // 2, 0, 0
    so.waveform, so.amp, so.w, so.tw, so.t0 = abi.WAVE_GAUSS_SINE, 1.0, 2 * np.pi, 0.3, 1.0
    o = Oracle(cfg)
    o.add_probe(abi.EZ, nx // 2 + 30, 0, 0)
    o.step(steps)
    return o.probes()[:, 0]


def test_split_without_sigma_is_the_plain_update():
    n, d = (20, 18, 16), 0.05
    base = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, 0.8))
    base.npml[:] = [4] * 6
    base.pml_R = 1.0  # ln(1/R) = 0: sigma = 0 everywhere
    rng = np.random.RandomState(2)
    init = [rng.standard_normal(n) for _ in range(6)]
    out = []
    for kind in (abi.PML_CPML, abi.PML_SPLIT):
        cfg = copy.deepcopy(base)
        cfg.pml_kind = kind
        o = Oracle(cfg)
        for c in range(6):
            o.set_field(c, init[c])
        o.step(30)
        out.append(o.fields())
    for a, b in zip(*out):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


def test_split_layer_absorbs():
    steps = 900
    short = pulse(140, steps, 16, abi.PML_SPLIT)
    long_ = pulse(600, steps, 16, abi.PML_SPLIT)
    refl = np.abs(short - long_).max() / np.abs(long_).max()
    assert refl < 5e-3


def test_split_and_cpml_agree():
    steps = 900
    s = pulse(140, steps, 16, abi.PML_SPLIT)
    c = pulse(140, steps, 16, abi.PML_CPML)
    none = pulse(140, steps, 0, abi.PML_CPML)
    scale = np.abs(c).max()
    assert np.abs(s - c).max() / scale < 1e-2
    assert np.abs(none - c).max() / scale > 5 * np.abs(s - c).max() / scale


def test_split_single_cell_closed_form():
    """One H component in a corner cell of an x/z layer after one step from
    a prescribed E: P1' = k1_a1 P1 - k2_a1 dt d1, P2' = k1_a2 P2 + k2_a2 dt d2."""
    n, d = (16, 12, 16), 0.05
    cfg = abi.default_config(*n, d, d, d, configs.dt_for(d, d, d, 0.8))
    cfg.npml[:] = [5, 5, 0, 0, 5, 5]
    cfg.pml_kind = abi.PML_SPLIT
    rng = np.random.RandomState(9)
    E = [rng.standard_normal(n) for _ in range(3)]
    H0 = rng.standard_normal(n)
    o = Oracle(cfg)
    for c in range(3):
        o.set_field(c, E[c])
    o.set_field(abi.HY, H0)      # all of it in P1 (the stored P2 starts at 0)
    o.step(1)
    i, j, k = 2, 6, 3            # inside the x and z slabs, stencil inside the grid
    # Hy (2i, 2j+1, 2k): first curl term D_z Ex (a1 = z), second D_x Ez (a2 = x)
    from oracle import oracle as orc
    def d24(f0, f1, f2, f3): return (27 * (f2 - f1) - (f3 - f0)) / (24 * d)
    Ex, Ez = E[0], E[2]
    dzEx = d24(Ex[i, j, k - 2], Ex[i, j, k - 1], Ex[i, j, k], Ex[i, j, k + 1])
    dxEz = d24(Ez[i - 2, j, k], Ez[i - 1, j, k], Ez[i, j, k], Ez[i + 1, j, k])
    sz = orc.pml_sigma(2 * k, n[2], 5, 5, d, cfg.pml_order, cfg.pml_R, cfg.pml_vmax)
    sx = orc.pml_sigma(2 * i, n[0], 5, 5, d, cfg.pml_order, cfg.pml_R, cfg.pml_vmax)
    k2z, k2x = 1 / (1 + 0.5 * cfg.dt * sz), 1 / (1 + 0.5 * cfg.dt * sx)
    assert sz > 0 and sx > 0
    p1 = (2 * k2z - 1) * H0[i, j, k] - k2z * cfg.dt * dzEx
    p2 = k2x * cfg.dt * dxEz
    assert o.field(abi.HY)[i, j, k] == pytest.approx(p1 + p2, rel=1e-12, abs=1e-12)
