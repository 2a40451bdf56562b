"""dtmaxwell4_amd — MI355X-native FDTD stepping path of DTmaxwell4.

The stepping path (H/E (2,4) sweeps with CPML, TF/SF, dispersive and
gyrotropic media, sensors and FP64 DFT) runs in hand-written gfx950 kernels
in libdtmaxwell4.so behind the C ABI of include/dtmaxwell4.h.  This package
holds the ctypes handle (`Solver`), the reference-compatible `DTmxw` module,
config helpers (`configs`) and the adjoint (`adjoint`).
"""
from . import abi  # noqa: F401


def __getattr__(name):
    # submodules through importlib: `from . import X` would consult this hook
    # again before X is imported
    import importlib
    if name == "Solver":
        return importlib.import_module(".solver", __name__).Solver
    if name == "DTmxw":
        return importlib.import_module(".DTmxw", __name__)
    raise AttributeError(name)
