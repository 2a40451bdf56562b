"""Snapshots and restart: dtm_write_arr writes the aivlib .arr layout of the
reference's SeismoDrops::drop (drop.cu:84-107), and dtm_save_state /
dtm_load_state resume a run bit for bit (fields, CPML psi, dispersive
currents, TF/SF clock, DFT bins, flux totals) under both schedules and both
storage precisions."""
import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs

pytestmark = pytest.mark.gpu


def read_arr(path):
    """aivlib .arr reader: int comment size, comment, dim, value size, the
    extents fastest first, then the values."""
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:4], dtype=np.int32)[0]
    off = 4 + head
    dim, size = np.frombuffer(raw[off:off + 8], dtype=np.int32)
    ext = np.frombuffer(raw[off + 8:off + 8 + 4 * dim], dtype=np.int32)
    data = np.frombuffer(raw[off + 8 + 4 * dim:], dtype=np.float32 if size == 4 else np.float64)
    return head, dim, size, ext, data.reshape(tuple(ext[::-1]))


def build(storage, sweep):
    from dtmaxwell4_amd.solver import Solver
    assert abi.load_library().dtm_device_count() > 0, "no HIP device visible"
    w = configs.drude_lorentz_purcell(n=40, npml=6, radius=6, gap=3, nfreq=4, steps=0, cells_per_lambda=16,
                                      storage=storage)
    w.cfg.sweep = sweep
    s = configs.apply(w, Solver(w.cfg))
    s.add_flux_box(8, 32, 8, 32, 8, 32)
    return s


def test_arr_snapshot_layout(tmp_path):
    s = build(abi.STORE_FP32, abi.SWEEP_TWOPASS)
    s.step(60)
    p = tmp_path / "ragEz_00060.arr"
    s.write_arr(abi.EZ, p)
    head, dim, size, ext, data = read_arr(p)
    raw = np.frombuffer(open(p, "rb").read()[:36], dtype=np.int32)
    assert list(raw) == [12, 0, 0, 0, 3, 4, 40, 40, 40]
    assert (head, dim, size) == (12, 3, 4)
    ez = s.field(abi.EZ)
    assert np.abs(ez).max() > 0
    np.testing.assert_array_equal(data, ez)
    s.close()


@pytest.mark.parametrize("storage", [abi.STORE_FP32, abi.STORE_FP16], ids=["fp32", "fp16"])
@pytest.mark.parametrize("sweep", [abi.SWEEP_TWOPASS, abi.SWEEP_FUSED], ids=["twopass", "fused"])
def test_save_and_resume_is_bitwise(tmp_path, storage, sweep):
    a = build(storage, sweep)
    a.step(45)   # an odd count: the fused step's buffers are swapped
    path = tmp_path / "state.bin"
    a.save_state(path)
    a.step(70)
    fa, da, ta = a.fields(), a.dft(), a.flux_total()
    a.close()
    b = build(storage, sweep)
    b.load_state(path)
    assert b.time_step == 45
    b.step(70)
    fb, db, tb = b.fields(), b.dft(), b.flux_total()
    for x, y in zip(fa, fb):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(ta, tb)
    assert np.abs(fa[abi.EZ]).max() > 0 and np.abs(da).max() > 0
    # a solver of another shape refuses the file
    from dtmaxwell4_amd.solver import Solver
    other = Solver(abi.default_config(20, 20, 20, 0.05, 0.05, 0.05, 0.01))
    with pytest.raises(abi.DTMError):
        other.load_state(path)
    other.close()
    b.close()
