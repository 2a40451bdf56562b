"""dtm_energy_box_noalloc (no device allocation: static scratch) against
dtm_energy_box (scratch the library allocates) through the C ABI: the same
value bit for bit on two calls in a row, the second on a smaller box of
another, smaller solver; and boxes beyond its static scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from dtmaxwell4_amd import abi, configs

pytestmark = pytest.mark.gpu


def bind():
    lib = abi.load_library()
    f = lib.dtm_energy_box_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.POINTER(C.c_double)]
    return lib, f


def stepped(nx, ny, nz, steps):
    from dtmaxwell4_amd.solver import Solver
    w = configs.plane_wave_sphere(nx=nx, ny=ny, nz=nz, npml=8, inset=4, radius=6, steps=steps, periods_w=1.0,
                                  periods_t0=2.0)
    return configs.apply(w, Solver(w.cfg))


def test_noalloc_energy_equals_allocating_entry_point():
    lib, noalloc = bind()
    assert lib.dtm_device_count() > 0, "no HIP device visible"
    cases = [((80, 56, 52), 180, (4, 76, 3, 53, 2, 50)),   # first call
             ((48, 40, 36), 120, (10, 30, 8, 20, 6, 30))]  # then a smaller input
    for shape, steps, box in cases:
        s = stepped(*shape, steps)
        s.step(steps)
        ref = C.c_double()
        abi.check(lib.dtm_energy_box(s.h, *box, C.byref(ref)), "dtm_energy_box")
        got = C.c_double(-1.0)
        abi.check(noalloc(s.h, *box, C.byref(got)), "dtm_energy_box_noalloc")
        print(shape, box, "allocating", ref.value, "static scratch", got.value)
        assert ref.value > 0
        assert got.value == ref.value
        s.close()


def test_noalloc_energy_refuses_boxes_beyond_its_scratch():
    from dtmaxwell4_amd.solver import Solver
    lib, noalloc = bind()
    cfg = abi.default_config(8200, 4, 4, 0.05, 0.05, 0.05, configs.dt_for(0.05, 0.05, 0.05))
    s = Solver(cfg)
    v = C.c_double()
    assert noalloc(s.h, 0, 8193, 0, 4, 0, 4, C.byref(v)) == -6  # DTM_ERR_LIMIT
    assert noalloc(s.h, 0, 8192, 0, 4, 0, 4, C.byref(v)) == 0
    assert v.value == 0.0
    s.close()
