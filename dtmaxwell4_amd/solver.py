"""Python handle over the explicit C ABI (dtm_*) of libdtmaxwell4.so.

    from dtmaxwell4_amd import Solver, abi
    cfg = abi.default_config(128, 128, 128)
    s = Solver(cfg)
    s.add_sensor("Ez", 64, 64, 64)
    s.step(100)
    series = s.probes()

Every call goes to the HIP library; a missing library raises (no fallback).
"""
import ctypes as C

import numpy as np

from . import abi


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Solver:
    """One FDTD grid on the current HIP device (dtm_create / dtm_destroy)."""

    def __init__(self, cfg):
        self.L = abi.load_library()
        self.cfg = cfg
        self.shape = (cfg.nx, cfg.ny, cfg.nz)
        h = C.c_void_p()
        abi.check(self.L.dtm_create(C.byref(cfg), C.byref(h)), "dtm_create")
        self.h = h
        self.nprobe = 0
        self.nfreq = 0
        self._design_shape = (0, 0, 0)  # no design box yet

    def close(self):
        if getattr(self, "h", None):
            self.L.dtm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- setup
    def set_materials(self, mx=None, my=None, mz=None):
        arrs = [None if m is None else np.ascontiguousarray(m, dtype=np.uint8) for m in (mx, my, mz)]
        for a in arrs:
            if a is not None and a.shape != self.shape:
                raise ValueError(f"material volume shape {a.shape} != {self.shape}")
        abi.check(self.L.dtm_set_materials(self.h, *[None if a is None else _ptr(a) for a in arrs]),
                  "dtm_set_materials")

    def set_design_eps(self, i0, j0, k0, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        ni, nj, nk = eps.shape
        abi.check(self.L.dtm_set_design_eps(self.h, i0, j0, k0, ni, nj, nk, _ptr(eps)), "dtm_set_design_eps")
        self._design_shape = (ni, nj, nk)

    def add_sensor(self, fld, x, y, z):
        r = abi.check(self.L.dtm_add_sensor(self.h, fld.encode(), x, y, z), "dtm_add_sensor")
        self.nprobe += 1
        return r

    def add_probe(self, comp, i, j, k):
        r = abi.check(self.L.dtm_add_probe(self.h, comp, i, j, k), "dtm_add_probe")
        self.nprobe += 1
        return r

    def set_dft(self, omegas):
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        abi.check(self.L.dtm_set_dft(self.h, len(om), _ptr(om)), "dtm_set_dft")
        self.nfreq = len(om)

    # ------------------------------------------------------------- stepping
    def step(self, n=1):
        abi.check(self.L.dtm_step(self.h, n), "dtm_step")

    def sync(self):
        abi.check(self.L.dtm_sync(self.h), "dtm_sync")

    def time_steps(self, n):
        ms = C.c_double()
        abi.check(self.L.dtm_time_steps(self.h, n, C.byref(ms)), "dtm_time_steps")
        return ms.value

    @property
    def time_step(self):
        return self.L.dtm_time_step(self.h)

    @property
    def kernel_name(self):
        return self.L.dtm_kernel_name(self.h).decode()

    def set_kernel_timing(self, on=True):
        abi.check(self.L.dtm_set_kernel_timing(self.h, int(on)), "dtm_set_kernel_timing")

    def kernel_times(self):
        """Per-group (H interior, H edge, E interior, E edge, aux) summed ms and
        launch counts since the last call."""
        ms = np.zeros(5, dtype=np.float64)
        cnt = np.zeros(5, dtype=np.int64)
        abi.check(self.L.dtm_kernel_times(self.h, _ptr(ms), _ptr(cnt)), "dtm_kernel_times")
        return ms, cnt

    def group_cells(self, g):
        return self.L.dtm_group_cells(self.h, g)

    # --------------------------------------------------------------- access
    def field(self, comp):
        out = np.empty(self.shape, dtype=np.float32)
        abi.check(self.L.dtm_field_read(self.h, comp, _ptr(out)), "dtm_field_read")
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError("field shape mismatch")
        abi.check(self.L.dtm_field_write(self.h, comp, _ptr(a)), "dtm_field_write")

    def fields(self):
        return [self.field(c) for c in range(6)]

    def probes(self):
        rows = self.time_step
        out = np.zeros((max(rows, 1), max(self.nprobe, 1)), dtype=np.float32)
        r = abi.check(self.L.dtm_probe_read(self.h, _ptr(out), rows), "dtm_probe_read")
        return out[:r, :self.nprobe]

    def probe_last(self):
        """Probe values after the last step (None before the first step)."""
        out = np.zeros(max(self.nprobe, 1), dtype=np.float32)
        r = abi.check(self.L.dtm_probe_last(self.h, _ptr(out)), "dtm_probe_last")
        return out[:self.nprobe] if r > 0 else None

    def dft(self):
        out = np.zeros((max(self.nprobe, 1), max(self.nfreq, 1), 2), dtype=np.float64)
        abi.check(self.L.dtm_dft_read(self.h, _ptr(out)), "dtm_dft_read")
        out = out[:self.nprobe, :self.nfreq]
        return out[..., 0] + 1j * out[..., 1]

    def energy(self, box=None):
        """1/2 sum (eps E^2 + H^2) dV in FP64 over the grid, or over the cells
        box = (i0, i1, j0, j1, k0, k1)."""
        v = C.c_double()
        if box is None:
            abi.check(self.L.dtm_energy(self.h, C.byref(v)), "dtm_energy")
        else:
            abi.check(self.L.dtm_energy_box(self.h, *[int(b) for b in box], C.byref(v)), "dtm_energy_box")
        return v.value

    def set_tfsf_fold(self, on=True):
        """Tabulated plane-wave TF/SF inside the sweeps (default) or in the
        separate correction kernel; same fields."""
        abi.check(self.L.dtm_set_tfsf_fold(self.h, int(bool(on))), "dtm_set_tfsf_fold")

    def set_fp32_zone(self, i0, i1, j0, j1, k0, k1):
        """FP16 storage: keep the cells [i0,i1) x [j0,j1) x [k0,k1) in FP32 too
        (around a dipole source and its probes); before the first step."""
        abi.check(self.L.dtm_set_fp32_zone(self.h, i0, i1, j0, j1, k0, k1), "dtm_set_fp32_zone")

    def add_flux_box(self, i0, i1, j0, j1, k0, k1):
        """Record the FP64 Poynting flux through the six faces of a box every
        step; returns the box id."""
        r = abi.check(self.L.dtm_add_flux_box(self.h, i0, i1, j0, j1, k0, k1), "dtm_add_flux_box")
        self.nflux = getattr(self, "nflux", 0) + 1
        return r

    def flux(self):
        """Flux series [step][box][face] (faces x-, x+, y-, y+, z-, z+, outward)."""
        nb = getattr(self, "nflux", 0)
        rows = self.time_step
        out = np.zeros((max(rows, 1), max(6 * nb, 1)), dtype=np.float64)
        r = abi.check(self.L.dtm_flux_read(self.h, _ptr(out), rows), "dtm_flux_read")
        return out[:r, :6 * nb].reshape(r, nb, 6)

    def flux_total(self):
        """Running FP64 sums of flux * dt: [box][face]."""
        nb = getattr(self, "nflux", 0)
        out = np.zeros(max(6 * nb, 1), dtype=np.float64)
        abi.check(self.L.dtm_flux_total(self.h, _ptr(out)), "dtm_flux_total")
        return out[:6 * nb].reshape(nb, 6)

    def write_arr(self, comp, path):
        """One component as an aivlib .arr snapshot (drop.cu:84-107 layout)."""
        abi.check(self.L.dtm_write_arr(self.h, comp, str(path).encode()), "dtm_write_arr")

    def save_state(self, path):
        """The whole time-stepped state, step count, DFT bins and flux totals."""
        abi.check(self.L.dtm_save_state(self.h, str(path).encode()), "dtm_save_state")

    def load_state(self, path):
        """Resume from dtm_save_state (same config, materials, probes, boxes)."""
        abi.check(self.L.dtm_load_state(self.h, str(path).encode()), "dtm_load_state")

    def set_adjoint_checkpoints(self, every):
        """Adjoint memory scheme: 0 keeps the design-box E of every step;
        K > 0 keeps the whole state every K steps and recomputes each segment."""
        abi.check(self.L.dtm_adjoint_set_checkpoints(self.h, int(every)), "dtm_adjoint_set_checkpoints")

    def design_gradient(self, nsteps, omega, monitors):
        """Adjoint gradient of J = sum_m |DFT_omega Ez_m|^2 over nsteps with
        respect to the design box's 1/eps.  monitors: (i, j, k) cells.
        Returns (J, dJ/d(1/eps) as an array shaped like the design box).
        The solver is left at step 0 with zero fields."""
        mon = np.ascontiguousarray(np.asarray(monitors, dtype=np.int32).reshape(-1, 3))
        shape = self._design_shape
        grad = np.zeros(shape, dtype=np.float64)
        J = C.c_double()
        abi.check(self.L.dtm_adjoint_design_gradient(self.h, int(nsteps), float(omega), mon.shape[0],
                                                      mon.ctypes.data, C.byref(J), grad.ctypes.data),
                  "dtm_adjoint_design_gradient")
        return J.value, grad
