"""The headline workload at its full size (TF/SF plane wave on an eps=4
sphere, 512x256x256, 12-cell CPML, TF/SF box 16 cells inside the CPML),
checked through properties that do not need the FP64 oracle at this size:
the two time-step schedules agree, the scattered-field region stays empty
without a scatterer, and FP16 storage tracks FP32.  The pulse is shortened
(1 period wide, 2 periods delay) so that 400 steps carry it onto the sphere."""
import numpy as np
import pytest

from conftest import rel_l2
from dtmaxwell4_amd import abi, configs

pytestmark = pytest.mark.gpu

STEPS = 400


def run(storage=abi.STORE_FP32, sweep=abi.SWEEP_TWOPASS, radius=40, eps=4.0):
    from dtmaxwell4_amd.solver import Solver
    assert abi.load_library().dtm_device_count() > 0, "no HIP device visible"
    w = configs.plane_wave_sphere(storage=storage, radius=radius, eps=eps, steps=STEPS, periods_w=1.0,
                                  periods_t0=2.0)
    w.cfg.sweep = sweep
    s = configs.apply(w, Solver(w.cfg))
    s.step(STEPS)
    f = s.fields()
    s.close()
    return f


def test_headline_schedules_agree():
    """Two-pass and fused steps compute the same update: FP32 fields agree to
    the FP32 parity bound after 400 steps at 512x256x256."""
    a = run(sweep=abi.SWEEP_TWOPASS)
    b = run(sweep=abi.SWEEP_FUSED)
    assert all(np.isfinite(x).all() for x in a)
    err = rel_l2(b, a)
    print(f"fused vs two-pass rel-L2 {err:.3e}")
    assert err <= 1e-5


def test_headline_no_scatterer_leaves_scattered_region_empty():
    """Without the sphere the total field stays inside the TF/SF box: the
    scattered-field slabs next to the x faces hold < 1% of the incident peak."""
    ez = run(radius=0, eps=1.0)[abi.EZ]
    lo = 12 + 16  # CPML + inset: the box's low x face
    inside = np.abs(ez[lo + 3:-lo - 3, lo + 3:-lo - 3, lo + 3:-lo - 3]).max()
    outside = max(np.abs(ez[12:lo - 3]).max(), np.abs(ez[-lo + 3:-12]).max())
    print("inside", inside, "outside", outside)
    assert inside > 0.1
    assert outside < 1e-2 * inside


def test_headline_fp16_storage_tracks_fp32():
    """FP16 storage (per-component power-of-two scale, FP32 arithmetic) stays
    within 2% relative L2 of FP32 storage at full size."""
    a = run()
    b = run(storage=abi.STORE_FP16)
    err = rel_l2(b, a)
    print(f"fp16 vs fp32 rel-L2 {err:.3e}")
    assert err < 2e-2


def test_fused_fold_with_capped_segments():
    """416 x 512 columns make 256 column tiles, as many as the blocks an
    MI355X holds at once, so the fused step's cost model would march each
    tile over all 300 x planes in one round; with the plane wave folded in,
    its per-block factor table caps segments at 261 planes, and every column
    is split in two (150 planes each).  The fused step (folded) must still
    match the two-pass step (folded)."""
    from dtmaxwell4_amd.solver import Solver
    out = []
    for sweep in (abi.SWEEP_TWOPASS, abi.SWEEP_FUSED):
        w = configs.plane_wave_sphere(nx=300, ny=416, nz=512, radius=40, steps=300, periods_w=1.0,
                                      periods_t0=2.0)
        w.cfg.sweep = sweep
        s = configs.apply(w, Solver(w.cfg))
        s.step(300)
        out.append(s.fields())
        s.close()
    assert all(np.isfinite(x).all() for x in out[1])
    assert np.abs(out[0][abi.EZ]).max() > 0.1
    err = rel_l2(out[1], out[0])
    print(f"fused vs two-pass rel-L2 {err:.3e}")
    assert err <= 1e-5
