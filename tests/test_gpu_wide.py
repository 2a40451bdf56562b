"""Grids of 2^32 cells and more (the sweeps' 64-bit-index variants).

A plane wave along -x in a grid periodic in y and z carries the same field in
every (y, z) column.  So a 16640 x 512 x 512 grid (4.36e9 cells; the cells of
its last 256 x planes lie beyond index 2^32) must give, at probes near its
high-x TF/SF face, in every column, the probes of a 16640 x 8 x 8 grid that
the 32-bit variants run: the same arithmetic per cell, so the bound is
1e-6 of the peak (measured 0).  A wrong 64-bit index would read another
column or the zero pad and break the equality.  FP32 and FP16 storage
(105 GB and 53 GB of fields).  The split-field layer and the FP32 zone have
no wide variants and are refused."""
import math

import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs

pytestmark = pytest.mark.gpu

NX, DX, B = 16640, 0.05, 16
STEPS = 300


def case(ny, nz, storage):
    cfg = abi.default_config(NX, ny, nz, DX, DX, DX, configs.dt_for(DX, DX, DX))
    cfg.npml[:] = [12, 12, 0, 0, 0, 0]
    cfg.bc[1] = cfg.bc[2] = abi.BC_PERIODIC
    cfg.storage = storage
    t = cfg.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_PLANE, abi.WAVE_GAUSS_SINE, 0
    t.box[:] = [B * DX, (NX - B) * DX, -1e9, 1e9, -1e9, 1e9]   # faces along x only
    t.w, t.cSrc, t.Imp, t.amp = 2 * math.pi, 1.0, 1.0, 1.0
    t.theta, t.phi, t.psi = math.pi / 2, math.pi, -math.pi / 2   # k = -x, E along z
    t.tw, t.t0, t.tStop = 1.0, 3.0, 1e30
    t.pos[:] = [(NX - B) * DX, 0.0, 0.0]
    return cfg


def run(ny, nz, storage, cols):
    from dtmaxwell4_amd.solver import Solver
    s = Solver(case(ny, nz, storage))
    i = NX - B - 24                      # 24 cells inside the TF face
    for j, k in cols:
        s.add_probe(abi.EZ, i, j, k)
        s.add_probe(abi.HY, i, j, k)
    s.step(STEPS)
    p = s.probes()
    s.close()
    return p, i


@pytest.mark.parametrize("storage", [abi.STORE_FP32, abi.STORE_FP16])
def test_wide_grid_equals_thin_grid(storage):
    cols = [(0, 0), (511, 511), (255, 128), (1, 510)]
    big, i = run(512, 512, storage, cols)
    assert (i * 512 + 511) * 512 + 511 > 2 ** 32 and (i * 512) * 512 > 2 ** 32
    thin, _ = run(8, 8, storage, [(0, 0)])
    peak = np.abs(thin).max()
    print("probe peak", peak, "index of the first probe", (i * 512) * 512)
    assert peak > 0.1
    for c in range(len(cols)):
        d = np.abs(big[:, 2 * c:2 * c + 2] - thin).max() / peak
        print("column", cols[c], "max |big - thin| / peak", d)
        assert d <= 1e-6


def test_wide_grid_refuses_the_split_layer():
    from dtmaxwell4_amd.solver import Solver
    cfg = case(512, 512, abi.STORE_FP16)
    cfg.pml_kind = abi.PML_SPLIT
    with pytest.raises(abi.DTMError):
        Solver(cfg)
