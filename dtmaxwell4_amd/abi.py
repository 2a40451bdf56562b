"""ctypes mirror of include/dtmaxwell4.h and the loader of libdtmaxwell4.so.

The structures here are byte-for-byte the C structs of the public header.
`load_library()` returns the in-tree HIP library and raises if it is missing:
there is no CPU fallback for the stepping path.
"""
import ctypes as C
import os

EX, EY, EZ, HX, HY, HZ = range(6)
COMP_NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
BC_ZERO, BC_PERIODIC = 0, 1
STORE_FP32, STORE_FP16 = 0, 1
SWEEP_TWOPASS, SWEEP_FUSED = 0, 1
PML_CPML, PML_SPLIT = 0, 1
PML_PROFILE_POLY, PML_PROFILE_REFERENCE = 0, 1
TFSF_NONE, TFSF_DIPOLE, TFSF_PLANE = 0, 1, 2
WAVE_RAISED_COS, WAVE_GAUSS_SINE, WAVE_GAUSS = 0, 1, 2
MAX_MATS, MAX_POLES, MAX_SOFT = 64, 4, 16

ERRORS = {
    0: "ok",
    -1: "invalid argument",
    -2: "Courant condition violated",
    -3: "device error",
    -4: "out of memory",
    -5: "invalid state",
    -6: "limit exceeded",
}


class LorenzDisp(C.Structure):
    """One Lorentz/Drude pole (dispersion.cpp:5-22)."""

    _fields_ = [("a0", C.c_double), ("a1", C.c_double), ("b0", C.c_double),
                ("b1", C.c_double), ("b2", C.c_double)]

    def set(self, Deps=1.0, wp=0.0, gp=0.0, gp_=0.0):
        """LorenzDisp::set (dispersion.cpp:8-14)."""
        self.a0, self.a1, self.b0, self.b1, self.b2 = Deps * wp * wp, Deps * gp_, wp * wp, 2 * gp, 1.0
        return self

    def setDrude(self, Deps=1.0, wp=0.0, gp=0.0, gp_=0.0):
        """LorenzDisp::setDrude (dispersion.cpp:15-21)."""
        self.a0, self.a1, self.b0, self.b1, self.b2 = Deps * wp * wp, Deps * gp_, 0.0, 2 * gp, 1.0
        return self


class TFSFsrc(C.Structure):
    """The reference's source block, field for field (py_consts.h:73-91)."""

    _fields_ = [("tStop", C.c_float), ("srcXs", C.c_float), ("srcXa", C.c_float),
                ("srcXv", C.c_float), ("sphR", C.c_float), ("BoxMs", C.c_float),
                ("BoxPs", C.c_float), ("BoxMa", C.c_float), ("BoxPa", C.c_float),
                ("BoxMv", C.c_float), ("BoxPv", C.c_float), ("V_max", C.c_float),
                ("start", C.c_int), ("wavelength", C.c_float), ("w", C.c_float),
                ("k", C.c_float), ("Omega", C.c_float), ("kEnv", C.c_float),
                ("theta", C.c_float), ("phi", C.c_float), ("cSrc", C.c_float),
                ("Imp", C.c_float), ("t0", C.c_float), ("tw", C.c_float), ("T", C.c_float)]


class CoffStruct(C.Structure):
    """Coefficient block of defs.h:88-118."""

    _fields_ = [(n, C.c_float) for n in (
        "deps", "depsXX", "depsYY", "depsZZ", "coffAnisoXY", "coffAnisoXZ", "coffAnisoYX",
        "coffAnisoYZ", "coffAnisoZX", "coffAnisoZY", "gyrationXX", "gyrationXY", "gyrationXZ",
        "gyrationYX", "gyrationYY", "gyrationYZ", "gyrationZX", "gyrationZY", "gyrationZZ")] + [
        ("isDisp", C.c_int)]


class Material(C.Structure):
    _fields_ = [("eps_inf", C.c_double), ("sigma", C.c_double), ("npoles", C.c_int),
                ("pole", LorenzDisp * MAX_POLES), ("gyro", C.c_int), ("gyro_wp", C.c_double),
                ("gyro_nu", C.c_double), ("gyro_wc", C.c_double * 3)]


class Tfsf(C.Structure):
    _fields_ = [("kind", C.c_int), ("waveform", C.c_int), ("sf_inside", C.c_int),
                ("box", C.c_double * 6), ("w", C.c_double), ("kEnv", C.c_double),
                ("cSrc", C.c_double), ("Imp", C.c_double), ("theta", C.c_double),
                ("phi", C.c_double), ("psi", C.c_double), ("t0", C.c_double),
                ("tw", C.c_double), ("amp", C.c_double), ("tStop", C.c_double),
                ("pos", C.c_double * 3)]


class SoftSource(C.Structure):
    _fields_ = [("comp", C.c_int), ("i", C.c_int), ("j", C.c_int), ("k", C.c_int),
                ("waveform", C.c_int), ("amp", C.c_double), ("w", C.c_double),
                ("t0", C.c_double), ("tw", C.c_double), ("kEnv", C.c_double)]


class Config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("bc", C.c_int * 3), ("npml", C.c_int * 6), ("pml_order", C.c_double),
                ("pml_R", C.c_double), ("pml_vmax", C.c_double), ("pml_kind", C.c_int),
                ("storage", C.c_int), ("scale_e_log2", C.c_int * 3),
                ("scale_h_log2", C.c_int * 3), ("nmat", C.c_int),
                ("mat", Material * MAX_MATS), ("tfsf", Tfsf), ("ntime", C.c_int),
                ("nsoft", C.c_int), ("soft", SoftSource * MAX_SOFT), ("sweep", C.c_int),
                ("pml_profile", C.c_int)]


def default_config(nx, ny, nz, dx=0.005, dy=0.005, dz=0.005, dt=0.002):
    """Same defaults as dtm_config_defaults (reference values of py_consts.h:71,
    DTmxw.cu:224-225)."""
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.dx, c.dy, c.dz, c.dt = dx, dy, dz, dt
    c.pml_order, c.pml_R, c.pml_vmax = 4.0, 1e-4, 1.0
    c.nmat = 1
    c.mat[0].eps_inf = 1.0
    c.tfsf.amp = 1.0
    c.tfsf.cSrc = 1.0
    c.tfsf.Imp = 1.0
    c.tfsf.sf_inside = 1
    c.ntime = 1
    return c


def courant(c):
    """Normalised Courant number dt*sqrt(sum 1/d^2) (DTmxw.cu:332 bound: 6/7)."""
    return c.dt * (1 / c.dx ** 2 + 1 / c.dy ** 2 + 1 / c.dz ** 2) ** 0.5


_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdtmaxwell4.so")
_lib = None


class DTMError(RuntimeError):
    pass


def load_library():
    """Load the in-tree HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DTMError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback for the FDTD stepping path)")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    P = C.c_void_p
    sig = {
        "dtm_last_error": (C.c_char_p, []),
        "dtm_version": (C.c_char_p, []),
        "dtm_device_count": (C.c_int, []),
        "dtm_config_defaults": (None, [C.POINTER(Config), C.c_int, C.c_int, C.c_int]),
        "dtm_create": (C.c_int, [C.POINTER(Config), C.POINTER(P)]),
        "dtm_destroy": (None, [P]),
        "dtm_set_materials": (C.c_int, [P, P, P, P]),
        "dtm_set_design_eps": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P]),
        "dtm_add_sensor": (C.c_int, [P, C.c_char_p, C.c_int, C.c_int, C.c_int]),
        "dtm_add_probe": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
        "dtm_set_dft": (C.c_int, [P, C.c_int, P]),
        "dtm_step": (C.c_int, [P, C.c_int]),
        "dtm_time_step": (C.c_longlong, [P]),
        "dtm_sync": (C.c_int, [P]),
        "dtm_probe_count": (C.c_int, [P]),
        "dtm_probe_read": (C.c_int, [P, P, C.c_int]),
        "dtm_probe_last": (C.c_int, [P, P]),
        "dtm_dft_read": (C.c_int, [P, P]),
        "dtm_field_read": (C.c_int, [P, C.c_int, P]),
        "dtm_field_write": (C.c_int, [P, C.c_int, P]),
        "dtm_energy": (C.c_int, [P, C.POINTER(C.c_double)]),
        "dtm_energy_box": (C.c_int, [P] + [C.c_int] * 6 + [C.POINTER(C.c_double)]),
        "dtm_add_flux_box": (C.c_int, [P] + [C.c_int] * 6),
        "dtm_set_fp32_zone": (C.c_int, [P] + [C.c_int] * 6),
        "dtm_set_tfsf_fold": (C.c_int, [P, C.c_int]),
        "dtm_flux_read": (C.c_int, [P, P, C.c_int]),
        "dtm_flux_total": (C.c_int, [P, P]),
        "dtm_write_arr": (C.c_int, [P, C.c_int, C.c_char_p]),
        "dtm_save_state": (C.c_int, [P, C.c_char_p]),
        "dtm_load_state": (C.c_int, [P, C.c_char_p]),
        "dtm_adjoint_design_gradient": (C.c_int, [P, C.c_int, C.c_double, C.c_int, P, C.POINTER(C.c_double), P]),
        "dtm_adjoint_set_checkpoints": (C.c_int, [P, C.c_int]),
        "dtm_time_steps": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
        "dtm_kernel_name": (C.c_char_p, [P]),
        "dtm_set_kernel_timing": (C.c_int, [P, C.c_int]),
        "dtm_kernel_times": (C.c_int, [P, P, P]),
        "dtm_group_cells": (C.c_longlong, [P, C.c_int]),
        "TFSFsrc_set": (None, [C.POINTER(TFSFsrc), C.c_double, C.c_double, C.c_double]),
        "TFSFsrc_check": (None, [C.POINTER(TFSFsrc)]),
        "CoffStruct_set_eps": (None, [C.POINTER(CoffStruct), C.c_double, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double]),
        "_main": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
        "calcStep": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc, what="dtm call"):
    if rc < 0:
        msg = _lib.dtm_last_error().decode() if _lib is not None else ""
        raise DTMError(f"{what} failed: {ERRORS.get(rc, rc)}: {msg}")
    return rc
