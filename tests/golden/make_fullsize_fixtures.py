"""Writes the FP64-oracle fixtures of the workloads at their full sizes and
step counts, where the oracle needs an hour or more and cannot run inside a
GPU test:

  plane_wave_sphere      512x256x256, 4000 steps (TF/SF plane wave on an
                         eps=4 sphere, CPML 12, box 16 cells inside it)
  drude_lorentz_purcell  384^3, 4000 steps, 64 FP64 DFT bins (Drude+Lorentz
                         sphere, dipole TF/SF 4 cells from its surface)
  gyro_thin              the gyrotropic slab on an 8x8x512 cross-section,
                         8000 steps (the workload is periodic and uniform in
                         x and y, so its 256x256x512 fields equal these)

Each fixture holds the probe series of every step, the DFT bins, and at
steps `snap_steps` a plane of all six components (every second cell along
its two axes, float32), per-component whole-grid sums of squares and, for
the plane wave, the scattered-field L2 and the total-field energy.  The
numbers come from the oracle (oracle/fdtd_oracle.c, FP64); the inputs are
the workload builders of dtmaxwell4_amd.configs with their defaults.

    OMP_NUM_THREADS=6 python tests/golden/make_fullsize_fixtures.py plane_wave_sphere
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from dtmaxwell4_amd import abi, configs  # noqa: E402


def workload(name):
    """The workload and the extra probes, plane axis and index of a fixture
    (the GPU tests build the same through this function)."""
    if name == "plane_wave_sphere":
        w = configs.plane_wave_sphere()
        nx, ny, nz = w.cfg.nx, w.cfg.ny, w.cfg.nz
        pts = configs.seeded_points(8, (nx, ny, nz), 14, (nx // 2, ny // 2, nz // 2), 0, seed=2024)
        extra = [(c, *p) for p in pts for c in (abi.EZ, abi.HY)]
        return w, extra, 2, nz // 2, (2000, 4000)
    if name == "drude_lorentz_purcell":
        w = configs.drude_lorentz_purcell()
        ci = w.notes["dipole"][0]
        return w, [], 0, ci, (2000, 4000)
    if name == "gyro_thin":
        w = configs.gyro_slab(nx=8, ny=8)
        return w, [], 0, 4, (4000, 8000)
    raise ValueError(name)


def plane(f, axis, idx):
    """Every second cell of the plane `idx` normal to `axis`."""
    sl = [slice(None, None, 2)] * 3
    sl[axis] = idx
    return np.ascontiguousarray(f[tuple(sl)])


def eps_volumes(w):
    out = []
    for c in range(3):
        table = np.array([w.cfg.mat[m].eps_inf for m in range(w.cfg.nmat)])
        m = w.mats[c]
        out.append(np.ones((w.cfg.nx, w.cfg.ny, w.cfg.nz)) if m is None else table[m])
    return out


def scattered_mask(w):
    """Cells outside the TF/SF box (total-field box inside) and outside the
    CPML: the scattered-field region of a plane-wave workload."""
    c = w.cfg
    n = (c.nx, c.ny, c.nz)
    d = (c.dx, c.dy, c.dz)
    inside = np.ones(n, dtype=bool)
    keep = np.ones(n, dtype=bool)
    for a in range(3):
        idx = np.arange(n[a])
        x = (idx + 0.5) * d[a]
        ok = (x > c.tfsf.box[2 * a]) & (x < c.tfsf.box[2 * a + 1])
        sh = [1, 1, 1]
        sh[a] = n[a]
        inside &= ok.reshape(sh)
        keep &= ((idx >= c.npml[2 * a]) & (idx < n[a] - c.npml[2 * a + 1])).reshape(sh)
    return keep & ~inside


def summaries(w, fields, name):
    out = dict(sumsq=np.array([float(np.sum(f.astype(np.float64) ** 2)) for f in fields]))
    if name == "plane_wave_sphere":
        m = scattered_mask(w)
        out["scat_l2"] = np.array(float(np.sqrt(sum(float(np.sum(f[m].astype(np.float64) ** 2)) for f in fields))))
        ev = eps_volumes(w)
        en = sum(float(np.sum(ev[c] * fields[c].astype(np.float64) ** 2)) for c in range(3))
        en += sum(float(np.sum(fields[c].astype(np.float64) ** 2)) for c in range(3, 6))
        out["energy"] = np.array(en * w.cfg.dx * w.cfg.dy * w.cfg.dz)
    return out


def main(name):
    from oracle.oracle import Oracle
    w, extra, axis, idx, snaps = workload(name)
    o = configs.apply(w, Oracle(w.cfg))
    for p in extra:
        o.add_probe(*p)
    res = {}
    done = 0
    t0 = time.time()
    for s in snaps:
        o.step(s - done)
        done = s
        f = o.fields()
        res[f"plane_{s}"] = np.stack([plane(x, axis, idx) for x in f]).astype(np.float32)
        for k, v in summaries(w, f, name).items():
            res[f"{k}_{s}"] = v
        print(f"{name}: step {s} after {time.time() - t0:.0f} s", flush=True)
    res["probes"] = o.probes()
    if w.omegas is not None and len(w.omegas):
        d = o.dft()
        res["dft"] = np.stack([d.real, d.imag], -1)
    res["snap_steps"] = np.array(snaps)
    res["plane_axis"] = np.array(axis)
    res["plane_index"] = np.array(idx)
    o.close()
    path = os.path.join(HERE, f"fullsize_{name}.npz")
    np.savez_compressed(path, **res)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1])
