"""The reference's own driver surface end to end: a run.py-style script sets
DTmxw.cvar (grid, Tsteps) and cvar.shotpoint as run.py:11-37 does, calls
DTmxw._main([..., "--test", "--drop_dir", d]) and then calcStep once more.

Checked against the reference's output contract and the FP64 oracle:
  * one rate line per calcStep in the format of diamond.cu:483-485;
  * the sensor files <fld>-x%05dy%05dz%05d.dat (sensor.h:17), one "%g" line
    per calcStep (sensor.h:81), for the default sensor of init_index
    (DTmxw.cu:524-542) and a --sensor point;
  * every line equals the oracle's sensor value (sensor.h:51-83
    interpolation) after that calcStep's Ntime steps, on the reference
    geometry get_mat_indev (defs.h:128-148, restated below in float32 as the
    reference evaluates it) with the materials of init_material
    (DTmxw.cu:604-642) and the DipoleTFSF source (scattered field inside
    the box, raised-cosine window, Ntime gate of window.hpp:125)."""
import math
import os
import re

import numpy as np
import pytest

from dtmaxwell4_amd import abi

pytestmark = pytest.mark.gpu

NX, NY, NZ = 150, 150, 100
RATE = re.compile(r"^Step (\d+) /node 0/ subnode 0/: Time +[0-9.]+ ms \|drop +([0-9.]+)% \|\|rate +[0-9.]+ "
                  r"GYee_cells/sec \|total grid (\d+)x(\d+)x(\d+)=(\d+) cells \| isTFSF=([01])$", re.M)


def mat_at(ix, iy, iz, nx, ny, npmlx, npmly, dx, dy, dz):
    """get_mat_indev (defs.h:128-148) on half-coordinate arrays, float32."""
    f = np.float32
    IndAg, IndOrg, IndITO, IndGlass = 11, 12, 13, 14
    x = (ix * 0.5 * f(dx)).astype(f)
    y = (iy * 0.5 * f(dy)).astype(f)
    z = (iz * 0.5 * f(dz)).astype(f)
    Xc, Yc = f(0.5 * nx * f(dx)), f(0.5 * ny * f(dy))
    mx, my = f(f(2 * npmlx + 100) * f(dx)) / f(2), f(f(2 * npmly + 100) * f(dy)) / f(2)
    out = np.full(np.broadcast(x, y, z).shape, IndGlass, dtype=np.uint8)
    period, rad, height, shift = f(0.4), f(0.1), f(0.1), f(1e-5)

    def roundf(v):  # C roundf: halves away from zero (np.round rounds them to even)
        a = np.abs(v)
        r = np.floor(a)
        r = r + (a - r >= f(0.5)).astype(f)
        return np.copysign(r, v).astype(f)

    xr = x - Xc - roundf((x - Xc) / period) * period
    yr = y - Yc - roundf((y - Yc) / period) * period
    z1 = z - (f(0.150) - shift)
    z2 = z1 - (f(0.180) + shift)
    z3 = z2 - (f(0.100) + shift)
    out = np.where(z3 < 0, IndITO, out)
    out = np.where(z2 < 0, IndOrg, out)
    hole = (z1 > -height) & (xr * xr + yr * yr < rad * rad)
    out = np.where(z1 < 0, np.where(hole, IndOrg, IndAg), out)
    edge = (x < mx) | (x > f(nx * f(dx)) - mx) | (y < my) | (y > f(ny * f(dy)) - my)
    return np.where(edge, IndOrg, out).astype(np.uint8)


def legacy_config(cv):
    """The dtm_config that _main builds from the globals (dtm_legacy.cpp)."""
    sp = cv.shotpoint
    c = abi.default_config(cv.GridNx, cv.GridNy, cv.GridNz, cv.ds, cv.da, cv.dv, cv.dt)
    c.npml[:] = [cv.Npmlx, cv.Npmlx, cv.Npmly, cv.Npmly, 0, cv.Npmlz]  # no layer at low z (DTmxw.cu:262)
    c.pml_vmax = sp.V_max if sp.V_max > 0 else 1.0
    c.pml_kind = abi.PML_SPLIT                      # the reference's layer and sigma profile
    c.pml_profile = abi.PML_PROFILE_REFERENCE
    c.ntime = cv.Ntime
    eps = [1.0, 5.0, 5.0, 4.0, 1e4, 1, 4, 9, 16, 25, 1, 1, 1.75 ** 2, 1.8 ** 2, 1.5 ** 2]
    c.nmat = 15
    for m, e in enumerate(eps):
        c.mat[m].eps_inf = e
    c.mat[2].eps_inf, c.mat[2].npoles = 1.0, 2        # IndGold (DTmxw.cu:625-630)
    c.mat[2].pole[0].setDrude(8.5, 4.9 * math.pi, 0.21, 1e-3)
    c.mat[2].pole[1].set(1.0, 0.5 * 2 * math.pi, 30 * 2 * math.pi, 340 * math.pi)
    c.mat[11].eps_inf, c.mat[11].npoles = 1.0, 1      # IndAg (DTmxw.cu:632-635)
    c.mat[11].pole[0].setDrude(1.0, 136.66 / 3.0, 2.69e13 / 2 / 3e14, 0)
    t = c.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_DIPOLE, abi.WAVE_RAISED_COS, 1
    t.box[:] = [sp.BoxMs, sp.BoxPs, sp.BoxMa, sp.BoxPa, sp.BoxMv, sp.BoxPv]
    t.w, t.kEnv, t.cSrc, t.Imp, t.t0, t.tw, t.amp, t.tStop = sp.w, sp.kEnv, sp.cSrc, sp.Imp, sp.t0, sp.tw, 1.0, sp.tStop
    return c


def test_main_and_calcstep_write_reference_outputs(tmp_path, capfd):
    from dtmaxwell4_amd import DTmxw
    from oracle.oracle import Oracle
    assert abi.load_library().dtm_device_count() > 0, "no HIP device visible"
    cv = DTmxw.cvar
    saved = {k: getattr(cv, k) for k in ("GridNx", "GridNy", "GridNz", "Tsteps", "DropData")}
    try:
        cv.GridNx, cv.GridNy, cv.GridNz = NX, NY, NZ
        dx, dy, dz, dt = cv.ds, cv.da, cv.dv, cv.dt
        # run.py:19-37
        center = [NX / 2 * dx, NY / 2 * dy, 0.150 + 0.08]
        SS = cv.shotpoint
        SS.srcXs, SS.srcXa, SS.srcXv = center
        SS.BoxMs, SS.BoxPs = SS.srcXs - 5 * dx, SS.srcXs + 5 * dx
        SS.BoxMa, SS.BoxPa = SS.srcXa - 5 * dy, SS.srcXa + 5 * dy
        SS.BoxMv, SS.BoxPv = SS.srcXv - 5 * dz, SS.srcXv + 5 * dz
        diag = math.sqrt((SS.BoxPs - SS.BoxMs) ** 2 + (SS.BoxPa - SS.BoxMa) ** 2 + (SS.BoxMv - SS.BoxPv) ** 2)
        cL = 1.0 / 1.75
        SS.set(cL, cL, 0.5 * cL)
        SS.tStop = diag / cL + 20 * 2 * math.pi / SS.w + SS.t0 + 10 * dt
        ntime = cv.Ntime
        cv.Tsteps = 3 * ntime
        # a sensor near the dipole, where the field has arrived after 3 Ntime steps
        extra = (int(round(center[0] / dx)) + 7, int(round(center[1] / dy)), int(round(center[2] / dz)) + 2)
        drop = str(tmp_path / "drop")
        # the DROP_DATA snapshots of Ez and Hx; a stale file of an earlier run goes
        cv.DropData = (1 << abi.EZ) | (1 << abi.HX)
        os.makedirs(drop)
        open(os.path.join(drop, "ragEy_00000.arr"), "wb").close()
        rc = DTmxw._main(["DTmxw", "--test", "--drop_dir", drop, "--sensor", "%d %d %d" % extra])
        assert rc == 0
        assert DTmxw.calcStep() == 0      # a driver may keep stepping after _main
        out = capfd.readouterr().out
        cfg = legacy_config(cv)
    finally:
        for k, v in saved.items():
            setattr(cv, k, v)

    lines = RATE.findall(out)
    assert [int(m[0]) for m in lines] == [0, 1, 2, 3], out[-2000:]
    for m in lines:
        assert tuple(int(v) for v in m[2:6]) == (NX, NY, NZ, NX * NY * NZ)
        assert m[6] == "1"
        assert 0.0 < float(m[1]) < 100.0  # the share of the snapshot writes (DropData on)
    # default sensor point of init_index and the --sensor point
    X0 = int(round((SS.BoxMs + SS.BoxPs) * 0.5 / dx))
    Y0 = int(round((SS.BoxMa + SS.BoxPa) * 0.5 / dy))
    Z0 = int(round(0.150 / dz))
    points = [(X0, Y0, Z0), extra]
    names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    series = {}
    for (x, y, z) in points:
        for f in names:
            path = os.path.join(drop, "%s-x%05dy%05dz%05d.dat" % (f, x, y, z))
            assert os.path.exists(path), path
            txt = open(path).read().split("\n")
            assert txt[-1] == "" and len(txt) == 5
            for v in txt[:-1]:
                assert v == "%g" % float(v)
            series[(f, x, y, z)] = np.array([float(v) for v in txt[:-1]])

    # the same run on the FP64 oracle, geometry restated from defs.h
    o = Oracle(cfg)
    off = [(0, 1, 1), (1, 0, 1), (1, 1, 0)]
    vol = []
    for c in range(3):
        ix = 2 * np.arange(NX)[:, None, None] + off[c][0]
        iy = 2 * np.arange(NY)[None, :, None] + off[c][1]
        iz = 2 * np.arange(NZ)[None, None, :] + off[c][2]
        vol.append(mat_at(ix, iy, iz, NX, NY, cfg.npml[0], cfg.npml[2], dx, dy, dz))
    assert {11, 12, 13, 14} <= set(np.unique(vol[2]).tolist())  # Ag, organic, ITO, glass all present
    o.set_materials(*vol)
    keys = list(series)
    for (f, x, y, z) in keys:
        o.add_sensor(f, x, y, z)
    ref = []
    for _ in range(4):
        o.step(ntime)
        ref.append([o.probe_now(1, names.index(f), x, y, z) for (f, x, y, z) in keys])
    ref = np.array(ref)
    of = o.fields()
    o.close()
    got = np.stack([series[k] for k in keys], 1)
    scale = np.abs(ref).max()
    assert scale > 0
    near = keys.index(("Ez",) + extra)
    assert np.abs(ref[:, near]).max() > 1e-3 * scale  # the --sensor point sees the dipole field
    err = np.abs(got - ref).max() / scale
    print("sensor files vs FP64 oracle: max |diff| / max |value|", err)
    # %g keeps 6 significant digits
    assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 2e-6 * scale)

    # DROP_DATA snapshots (drop.cu:75-107): one per calcStep and dropped
    # component, the aivlib header then x-major FP32 values
    assert not os.path.exists(os.path.join(drop, "ragEy_00000.arr"))
    for c, name in ((abi.EZ, "Ez"), (abi.HX, "Hx")):
        for it in range(4):
            path = os.path.join(drop, "rag%s_%05d.arr" % (name, it))
            raw = np.fromfile(path, dtype=np.int32, count=9)
            assert raw.tolist() == [12, 0, 0, 0, 3, 4, NZ, NY, NX], raw
            assert os.path.getsize(path) == 36 + 4 * NX * NY * NZ
        last = np.fromfile(path, dtype=np.float32, offset=36).reshape(NX, NY, NZ)
        e = np.linalg.norm(last - of[c]) / np.linalg.norm(of[c])
        print("rag%s_00003.arr vs FP64 oracle rel-L2" % name, e)
        assert e <= 1e-5
    assert not os.path.exists(os.path.join(drop, "ragEz_00004.arr"))
