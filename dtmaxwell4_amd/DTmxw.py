"""`DTmxw` — the reference's SWIG module surface (mxw.i:1-36) over ctypes.

A driver written against the reference's module runs unchanged:

    import DTmxw                       # or: from dtmaxwell4_amd import DTmxw
    SS = DTmxw.cvar.shotpoint
    SS.BoxMs, SS.BoxPs = ...
    SS.set(cL, cL, 0.5*cL)             # TFSFsrc::set (signal.hpp:99-115)
    DTmxw.cvar.Tsteps = 1000
    DTmxw._main(sys.argv)              # headless with --test

`cvar` exposes the library's globals through ctypes `in_dll`: Tsteps,
shotpoint and the grid constants of py_consts.h (GridNx, GridNy, GridNz, ds,
da, dv, dt, NDT, Npmlx, Npmly, Npmlz, Ntime, NyBloch, Nzw, Np).  Unlike the
reference, the grid globals are writable (no recompile per size).
"""
import ctypes as C

from . import abi

_L = abi.load_library()


def _set(self, LightSpeed, impedance, lmb):
    _L.TFSFsrc_set(C.byref(self), LightSpeed, impedance, lmb)


def _check(self):
    _L.TFSFsrc_check(C.byref(self))


abi.TFSFsrc.set = _set
abi.TFSFsrc.check = _check
TFSFsrc = abi.TFSFsrc

_INT = ("Tsteps", "DropData", "GridNx", "GridNy", "GridNz", "Npmlx", "Npmly", "Npmlz", "Ntime", "NyBloch", "Nzw", "Np")
_FLOAT = ("ds", "da", "dv", "dt")
_CONST = ("NDT",)


class _CVar:
    """SWIG-style `cvar`: attribute access to the library's C globals."""

    def __getattr__(self, name):
        if name == "shotpoint":
            return abi.TFSFsrc.in_dll(_L, "shotpoint")
        if name in _INT or name in _CONST:
            return C.c_int.in_dll(_L, name).value
        if name in _FLOAT:
            return C.c_float.in_dll(_L, name).value
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _CONST:
            raise AttributeError(f"{name} is read-only")
        if name in _INT:
            C.c_int.in_dll(_L, name).value = int(value)
        elif name in _FLOAT:
            C.c_float.in_dll(_L, name).value = float(value)
        elif name == "shotpoint":
            C.memmove(C.addressof(abi.TFSFsrc.in_dll(_L, "shotpoint")), C.addressof(value),
                      C.sizeof(abi.TFSFsrc))
        else:
            raise AttributeError(name)


cvar = _CVar()


def _main(argv):
    """int _main(int argc, char** argv) (DTmxw.cu:731); argv must be a list of str
    (the typemap of mxw.i:11-32)."""
    if not isinstance(argv, list):
        raise ValueError("Expecting a list")
    if not all(isinstance(a, str) for a in argv):
        raise ValueError("List items must be strings")
    arr = (C.c_char_p * (len(argv) + 1))(*[a.encode() for a in argv], None)
    return _L._main(len(argv), arr)


def calcStep():
    """int calcStep() (diamond.cu:335): advance Ntime steps."""
    return _L.calcStep()
