"""The C ABI library loads and exports every symbol include/dtmaxwell4.h
declares; the ctypes mirror matches the C struct layout.  CPU only: no
compute call is made."""
import ctypes as C
import os
import re
import subprocess

import pytest

from dtmaxwell4_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "dtmaxwell4.h")


def declared_symbols():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                           text, flags=re.M))
    funcs -= {"if", "while", "for", "return", "sizeof"}
    globs = set()
    for m in re.finditer(r"^\s*extern\s+(?:const\s+)?[A-Za-z_]+\s+([^;]+);", text, flags=re.M):
        for name in m.group(1).split(","):
            globs.add(name.strip())
    return funcs, globs


def test_library_exports_declared_symbols():
    lib = abi.load_library()
    funcs, globs = declared_symbols()
    assert "dtm_create" in funcs and "_main" in funcs and "calcStep" in funcs
    assert {"Tsteps", "shotpoint", "GridNx", "NDT"} <= globs
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted((funcs | globs) - exported)
    assert not missing, f"declared but not exported: {missing}"
    for f in funcs:
        getattr(lib, f)


PROBE = r'''
#include <stddef.h>
#include <stdio.h>
#include "dtmaxwell4.h"
int main(void){
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dtm_config), offsetof(dtm_config, mat),
         offsetof(dtm_config, tfsf), offsetof(dtm_config, soft), sizeof(TFSFsrc), sizeof(CoffStruct));
  printf("%zu %zu %zu\n", sizeof(dtm_material), sizeof(dtm_tfsf), sizeof(dtm_soft_source));
  return 0; }
'''


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    a = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert a[0] == C.sizeof(abi.Config)
    assert a[1] == abi.Config.mat.offset
    assert a[2] == abi.Config.tfsf.offset
    assert a[3] == abi.Config.soft.offset
    assert a[4] == C.sizeof(abi.TFSFsrc)
    assert a[5] == C.sizeof(abi.CoffStruct)
    assert a[6:] == [C.sizeof(abi.Material), C.sizeof(abi.Tfsf), C.sizeof(abi.SoftSource)]


def test_legacy_globals_have_reference_defaults():
    lib = abi.load_library()
    assert C.c_int.in_dll(lib, "NDT").value == 3            # py_consts.h:1
    assert abs(C.c_float.in_dll(lib, "dt").value - 0.002) < 1e-9  # py_consts.h:71
    assert abs(C.c_float.in_dll(lib, "ds").value - 0.005) < 1e-9
    assert C.c_int.in_dll(lib, "Tsteps").value == 220        # DTmxw.cu:730 (Ntime*10)
    assert C.c_int.in_dll(lib, "DropData").value == 0        # py_consts.h:19: DROP_DATA off by default


def test_tfsfsrc_set_matches_reference_formulas():
    """TFSFsrc::set (signal.hpp:99-115) through the C ABI."""
    import math
    lib = abi.load_library()
    s = abi.TFSFsrc()
    cL = 1.0 / 1.75
    lib.TFSFsrc_set(C.byref(s), cL, cL, 0.5 * cL)
    k = 2 * math.pi / (0.5 * cL)
    assert s.k == pytest.approx(k, rel=1e-6)
    assert s.w == pytest.approx(cL * k, rel=1e-6)
    assert s.kEnv == pytest.approx(0.05, rel=1e-6)
    assert s.theta == pytest.approx(math.pi, rel=1e-6)
    assert s.phi == pytest.approx(math.pi / 2, rel=1e-6)
    assert (s.t0, s.tw, s.T) == pytest.approx((0.5, 0.1, 2.5), rel=1e-6)
    assert s.V_max == pytest.approx(cL) and s.Imp == pytest.approx(cL) and s.cSrc == pytest.approx(cL)


def test_create_rejects_courant_violation_without_device():
    """dtm_create validates before touching the device (DTmxw.cu:332)."""
    lib = abi.load_library()
    cfg = abi.default_config(8, 8, 8, 0.005, 0.005, 0.005, 0.0025)  # 0.0025*sqrt(3)/0.005 = 0.866 > 6/7
    h = C.c_void_p()
    assert lib.dtm_create(C.byref(cfg), C.byref(h)) == -2
    assert b"Courant" in lib.dtm_last_error()


def test_coffstruct_set_eps_cayley_identities():
    """CoffStruct::set_eps (defs.h:96-117): deps*(2eps+G) = 2I and the
    gyration matrix is the Cayley transform (eps+G/2)^-1 (eps-G/2) -- checked
    on the C ABI with G = [[0,-gz,gy],[gz,0,-gx],[-gy,gx,0]] * w*dt."""
    import numpy as np
    lib = abi.load_library()
    c = abi.CoffStruct()
    eps, g, w, dt = 5.0, (0.0, 0.0, 0.1), 3.0, 0.002
    lib.CoffStruct_set_eps(C.byref(c), eps, 0, g[0], g[1], g[2], w, dt)
    gx, gy, gz = [v * w * dt for v in g]
    G = np.array([[0, -gz, gy], [gz, 0, -gx], [-gy, gx, 0]])
    D = np.array([[c.depsXX, c.coffAnisoXY, c.coffAnisoXZ], [c.coffAnisoYX, c.depsYY, c.coffAnisoYZ],
                  [c.coffAnisoZX, c.coffAnisoZY, c.depsZZ]])
    Gy = np.array([[c.gyrationXX, c.gyrationXY, c.gyrationXZ], [c.gyrationYX, c.gyrationYY, c.gyrationYZ],
                   [c.gyrationZX, c.gyrationZY, c.gyrationZZ]])
    A = 2 * eps * np.eye(3) + G
    np.testing.assert_allclose(D @ A, 2 * np.eye(3), atol=1e-6)
    np.testing.assert_allclose(Gy, np.linalg.solve(eps * np.eye(3) + G / 2, eps * np.eye(3) - G / 2), atol=1e-6)
