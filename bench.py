"""Headline benchmark of the FDTD stepping path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-fp16]

Workload (the reference metric's configuration): TF/SF plane wave
(x-propagating, Ez-polarised, Gaussian pulse) on a dielectric sphere
(radius 40 cells, eps = 4) in a 512x256x256 grid with a 12-cell CPML on all
faces and the TF/SF box 16 cells inside the CPML.  A step is one full
H + E update of the whole grid with every boundary, source and material term.
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  `value` = FP32-storage Gcell-updates/s; the same line
carries the FP16-storage rate, the dominant kernel's HBM roofline (algorithmic
bytes / average launch time from HIP events on the solver's stream) and the
CPU baseline (the FP32 OpenMP restatement of the same equations, timed on
this host's cores on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GROUPS_TWOPASS = ("H sweep", "-", "E sweep", "-", "TF/SF+ADE+sources+probes")
GROUPS_FUSED = ("fused step", "-", "-", "-", "TF/SF+ADE+sources+probes")
# Algorithmic bytes per cell, the roofline basis: two-pass half-step sweep:
# read 3E+3H, write 3 components (+1 class byte for E): 36 / 37 B in FP32.
# Fused step (H and E in one pass): read 3E+3H, write 3E+3H, read the class
# byte: 49 B.  The CPML psi (two FP32 values read and written per half-step
# for each slab a cell lies in: 16 B; 32 B per fused step) is region-local
# and reported beside it (psi_bytes, roofline.extra_bytes), not in it.
WORDS_PER_HALF = 9
WORDS_FUSED = 12
PSI_BYTES_HALF = 2 * 2 * 4


def psi_cells(cfg):
    """Sum over axes of the cells inside that axis' CPML slabs."""
    n = (cfg.nx, cfg.ny, cfg.nz)
    t = 0
    for a in range(3):
        o = n[(a + 1) % 3] * n[(a + 2) % 3]
        t += (cfg.npml[2 * a] + cfg.npml[2 * a + 1]) * o
    return t


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(kernel_prefix, sweep):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a plain copy; the file names
    its command): read MB + write MB of the first row whose kernel name starts
    with kernel_prefix, or None."""
    path = os.path.join(ROOT, "profiles", f"k_{'torre_fused' if sweep else 'update_twopass'}_fp32_512x256x256.txt")
    try:
        for line in open(path):
            if line.startswith(kernel_prefix):
                f = line[70:].split()
                return dict(bytes=(float(f[2]) + float(f[3])) * 1e6, source=os.path.relpath(path, ROOT))
    except OSError:
        return None
    return None


def gpu_run(storage, sweep, steps, warmup, nx, ny, nz):
    import numpy as np

    from dtmaxwell4_amd import abi, configs
    from dtmaxwell4_amd.solver import Solver
    w = configs.plane_wave_sphere(nx=nx, ny=ny, nz=nz, storage=storage)
    w.cfg.sweep = sweep
    s = configs.apply(w, Solver(w.cfg))
    log(f"[bench] {w.name} {nx}x{ny}x{nz} storage={'fp16' if storage else 'fp32'} "
        f"sweep={'fused' if sweep else 'twopass'} warmup={warmup} steps={steps}")
    s.step(warmup)
    s.sync()
    # throughput: HIP events around the K steps only (per-launch events would
    # add a gap after every kernel)
    t0 = time.perf_counter()
    ms = s.time_steps(steps)
    wall = time.perf_counter() - t0
    # per-kernel times for the roofline: a second pass with HIP events around
    # every launch group, on the solver's stream
    s.set_kernel_timing(True)
    s.kernel_times()  # reset
    s.time_steps(min(steps, 100))
    gms, gcnt = s.kernel_times()
    s.set_kernel_timing(False)
    cells = nx * ny * nz
    val = cells * steps / (ms * 1e-3) / 1e9
    # finite-field guard: the timed run must still be physical
    ez = s.field(abi.EZ)
    if not np.isfinite(ez).all():
        raise RuntimeError("non-finite field after timed run")
    esz = 2 if storage else 4
    fused = s.kernel_name.startswith("k_torre")
    names = GROUPS_FUSED if fused else GROUPS_TWOPASS
    pfrac = psi_cells(w.cfg) / cells  # slab memberships per cell
    groups = {}
    for g in range(5):
        if gcnt[g] == 0:
            continue
        avg = gms[g] / gcnt[g]
        gc = s.group_cells(g)
        # Yee bytes (the roofline basis) and the CPML psi bytes, reported apart
        if fused:
            byts = gc * (WORDS_FUSED * esz + 1)
            psib = gc * 2 * PSI_BYTES_HALF * pfrac
        else:
            byts = gc * (WORDS_PER_HALF * esz + (1 if g in (2, 3) else 0))
            psib = gc * PSI_BYTES_HALF * pfrac
        groups[names[g]] = dict(avg_ms=avg, launches=int(gcnt[g]), cells=int(gc),
                                algorithmic_bytes=int(byts), psi_bytes=int(psib),
                                gbs=(byts / (avg * 1e-3) / 1e9) if (gc and avg > 0) else None,
                                gbs_with_psi=((byts + psib) / (avg * 1e-3) / 1e9) if (gc and avg > 0) else None)
    dom = max((g for g in groups if groups[g]["cells"]), key=lambda g: groups[g]["avg_ms"] * groups[g]["launches"])
    bpc = (WORDS_FUSED * esz + 1) if fused else (2 * WORDS_PER_HALF * esz + 1)
    res = dict(gcells=val, ms=ms, wall_s=wall, groups=groups, dominant=dom, kernel=s.kernel_name,
               bytes_per_cell_update=bpc, fused=fused)
    s.close()
    return res


def adjoint_run(steps, every=200):
    """Adjoint/forward time ratio on the inverse-design workload: 400x200x200,
    12-cell CPML, 64x64x32 design box (eps = 1 + 11 rho, rho ~ U[0,1] seeded,
    3-cell box smoothed), soft Ez dipole, J = sum |DFT Ez|^2 over a 16x16
    monitor plane at one frequency.  Forward = plain stepping of the same
    steps; adjoint = dtm_adjoint_design_gradient (its recorded forward pass
    plus the backward pass), both wall-clock with the device synchronised;
    timed with the design-box history and with whole-state checkpoints every
    `every` steps plus recompute (the two gradients must be equal)."""
    import numpy as np

    from dtmaxwell4_amd import configs
    from dtmaxwell4_amd.solver import Solver
    w = configs.inverse_design(steps=steps)
    i0, j0, k0, eps = w.design
    mons = [(p[1], p[2], p[3]) for p in w.probes]
    omega = float(w.omegas[0])
    s = Solver(w.cfg)
    s.set_design_eps(i0, j0, k0, eps)
    log(f"[bench] adjoint workload {w.cfg.nx}x{w.cfg.ny}x{w.cfg.nz}, design {eps.shape}, {len(mons)} monitors, "
        f"{steps} steps")
    s.step(20)
    s.sync()
    t = time.perf_counter()
    s.step(steps)
    s.sync()
    fwd = time.perf_counter() - t
    # back to step 0: a fresh solver with the same set-up
    s.close()
    s = Solver(w.cfg)
    s.set_design_eps(i0, j0, k0, eps)
    t = time.perf_counter()
    J, grad = s.design_gradient(steps, omega, mons)
    adj = time.perf_counter() - t
    # the same gradient with whole-state checkpoints every `every` steps and recompute
    s.set_adjoint_checkpoints(every)
    t = time.perf_counter()
    Jc, gradc = s.design_gradient(steps, omega, mons)
    adjc = time.perf_counter() - t
    s.close()
    if not (np.isfinite(J) and np.isfinite(grad).all()):
        raise RuntimeError("non-finite adjoint result")
    if Jc != J or not np.array_equal(gradc, grad):
        raise RuntimeError("checkpointed adjoint differs from the history-mode adjoint")
    cells = w.cfg.nx * w.cfg.ny * w.cfg.nz
    return dict(ratio=adj / fwd, forward_s=fwd, adjoint_call_s=adj, steps=steps,
                checkpoint_every=every, ratio_checkpointed=adjc / fwd, adjoint_call_checkpointed_s=adjc,
                forward_gcells=cells * steps / fwd / 1e9, J=J, grad_norm=float(np.linalg.norm(grad)),
                workload=f"inverse design {w.cfg.nx}x{w.cfg.ny}x{w.cfg.nz}, design {list(eps.shape)}, "
                         f"{len(mons)} monitors, {steps} steps (call = recorded forward + backward; "
                         f"checkpointed: state every {every} steps + recompute)")


def purcell_run(steps, zone_margin=8):
    """Purcell error of FP16 storage on the dispersive dipole workload:
    384^3, 12-cell CPML, Drude (eps_inf 3.7, wp 1.37e16, gamma 2.7e13 rad/s) +
    Lorentz sphere r=40, Ez DipoleTFSF 4 cells above the surface, 64 FP64 DFT
    bins on 0.5-1.5 f0.  Error = max_w |P16 - P32| / max_w |P32| (purcell.gp:2).
    FP16 runs twice: plain, and with an FP32 zone of the cells within
    box_half + zone_margin of the dipole (dtm_set_fp32_zone)."""
    import numpy as np

    from dtmaxwell4_amd import abi, configs, purcell
    from dtmaxwell4_amd.solver import Solver
    out = {}
    curves = {}
    for name, store in (("fp32", abi.STORE_FP32), ("fp16", abi.STORE_FP16), ("fp16_zone", abi.STORE_FP16)):
        w = configs.drude_lorentz_purcell(steps=steps, storage=store)
        s = Solver(w.cfg)
        if name == "fp16_zone":
            s.set_fp32_zone(*configs.dipole_zone(w, zone_margin))
            out["fp32_zone"] = list(configs.dipole_zone(w, zone_margin))
        configs.apply(w, s)
        t = time.perf_counter()
        s.step(steps)
        s.sync()
        el = time.perf_counter() - t
        FE = s.dft()[0]
        s.close()
        Fd = purcell.dipole_spectrum(w.cfg.tfsf, w.omegas, steps, w.cfg.dt)
        curves[name] = purcell.purcell_factor(w.omegas, FE, Fd)
        cells = w.cfg.nx * w.cfg.ny * w.cfg.nz
        out[name + "_gcells"] = cells * steps / el / 1e9
        log(f"[bench] purcell {name}: {out[name + '_gcells']:.2f} Gcells/s, P in "
            f"[{curves[name].min():.4g}, {curves[name].max():.4g}]")
    P32, P16, P16z = curves["fp32"], curves["fp16"], curves["fp16_zone"]
    out["fp16_vs_fp32_error"] = float(np.max(np.abs(P16 - P32)) / np.max(np.abs(P32)))
    out["fp16_zone_vs_fp32_error"] = float(np.max(np.abs(P16z - P32)) / np.max(np.abs(P32)))
    out["P_fp32_range"] = [float(P32.min()), float(P32.max())]
    # against the FP64 oracle's spectrum of the same workload (fixture written
    # by tests/golden/make_fullsize_fixtures.py: the oracle needs hours here)
    fx = os.path.join(ROOT, "tests", "golden", "fullsize_drude_lorentz_purcell.npz")
    if steps == 4000 and os.path.exists(fx):
        d = np.load(fx)["dft"]
        w = configs.drude_lorentz_purcell(steps=steps)
        Fd = purcell.dipole_spectrum(w.cfg.tfsf, w.omegas, steps, w.cfg.dt)
        P64 = purcell.purcell_factor(w.omegas, d[0, :, 0] + 1j * d[0, :, 1], Fd)
        for name, P in (("fp32", P32), ("fp16", P16), ("fp16_zone", P16z)):
            out[f"{name}_vs_fp64_error"] = float(np.max(np.abs(P - P64)) / np.max(np.abs(P64)))
        out["fp64_reference"] = os.path.relpath(fx, ROOT)
    out["workload"] = (f"Drude+Lorentz sphere r=40 in 384^3, dipole TF/SF 4 cells from the surface, "
                       f"{len(P32)} DFT bins, {steps} steps")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(nx, ny, nz, budget_s=15.0):
    """FP32 OpenMP restatement of the same equations (oracle/liboracle32.so) on
    a bounded sample: the full grid for as many steps as fit ~budget_s."""
    from dtmaxwell4_amd import configs
    from oracle.oracle import Oracle, lib
    w = configs.plane_wave_sphere(nx=nx, ny=ny, nz=nz)
    o = configs.apply(w, Oracle(w.cfg, fp32=True))
    threads = lib(True).orc_nthreads()
    o.step(1)  # first touch of the arrays
    t = time.perf_counter()
    o.step(2)
    per = (time.perf_counter() - t) / 2
    n = max(1, min(400, int(budget_s / max(per, 1e-3))))
    t = time.perf_counter()
    o.step(n)
    el = time.perf_counter() - t
    o.close()
    cells = nx * ny * nz
    return dict(value=cells * n / el / 1e9, unit="Gcells/s", cores=threads, kind="port", cpu_model=cpu_model(),
                host_cpus=os.cpu_count(),
                sample=f"plane_wave_sphere {nx}x{ny}x{nz}, {n} full steps after 3 untimed steps "
                       f"({el:.1f} s), FP32 OpenMP restatement of the same update equations")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--size", type=str, default="512x256x256")
    ap.add_argument("--sweep", choices=("twopass", "fused"), default="twopass",
                    help="time-step schedule of the headline value; the other one is reported beside it")
    ap.add_argument("--no-alt", action="store_true", help="skip the other schedule")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward ratio run")
    ap.add_argument("--adjoint-steps", type=int, default=6000)
    ap.add_argument("--no-purcell", action="store_true", help="skip the FP16 Purcell-error run")
    ap.add_argument("--purcell-steps", type=int, default=4000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the benchmark, run a plain device copy of six grid-sized FP32 arrays "
                         "(the scale reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab sharding is a design note only).", file=sys.stderr)
        sys.exit(2)
    nx, ny, nz = [int(v) for v in args.size.split("x")]
    from dtmaxwell4_amd import abi
    sw = abi.SWEEP_FUSED if args.sweep == "fused" else abi.SWEEP_TWOPASS
    r32 = gpu_run(abi.STORE_FP32, sw, args.steps, args.warmup, nx, ny, nz)
    log(f"[bench] fp32 {r32['gcells']:.2f} Gcells/s  groups={json.dumps(r32['groups'])}")
    ralt = None
    if not args.no_alt:
        ralt = gpu_run(abi.STORE_FP32, 1 - sw, args.steps, args.warmup, nx, ny, nz)
        log(f"[bench] fp32 other schedule {ralt['gcells']:.2f} Gcells/s  groups={json.dumps(ralt['groups'])}")
    r16 = None
    if not args.no_fp16:
        r16 = gpu_run(abi.STORE_FP16, sw, args.steps, args.warmup, nx, ny, nz)
        log(f"[bench] fp16 {r16['gcells']:.2f} Gcells/s  groups={json.dumps(r16['groups'])}")
    adj = None
    if not args.no_adjoint:
        adj = adjoint_run(args.adjoint_steps)
        log(f"[bench] adjoint/forward {adj['ratio']:.3f} (forward {adj['forward_s']:.3f} s, "
            f"adjoint call {adj['adjoint_call_s']:.3f} s); checkpointed every {adj['checkpoint_every']}: "
            f"{adj['ratio_checkpointed']:.3f}")
    pur = None
    if not args.no_purcell:
        pur = purcell_run(args.purcell_steps)
        log(f"[bench] purcell fp16 vs fp32 error {pur['fp16_vs_fp32_error']:.3e} (with the FP32 zone "
            f"{pur['fp16_zone_vs_fp32_error']:.3e}), vs fp64 {pur.get('fp16_vs_fp64_error')} "
            f"({pur.get('fp16_zone_vs_fp64_error')})")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(nx, ny, nz)
            log(f"[bench] cpu baseline {cpu['value']:.4f} Gcells/s on {cpu['cores']} threads")
        except Exception as e:  # the baseline is reported, never the measured value
            log(f"[bench] cpu baseline failed: {e}")
    dom = r32["groups"][r32["dominant"]]
    achieved = dom["gbs"]
    # measured HBM bytes of the dominant kernel, from the committed profile
    prefix = {"E sweep": "void dtm::k_update_e<0, 3, 1,", "H sweep": "void dtm::k_update_h<0, 3, 1,",
              "fused step": "void dtm::k_torre2<0, 3>"}.get(r32["dominant"], "")
    tr = profiled_traffic(prefix, sw) if prefix and (nx, ny, nz) == (512, 256, 256) else None
    out = {
        "metric": "Gcells/s at 512x256x256 FP32 vs FP16-store; adjoint/forward time ratio",
        "value": round(r32["gcells"], 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"TF/SF plane wave on eps=4 sphere r=40, {nx}x{ny}x{nz}, CPML 12, "
                               "TF/SF box 16 cells inside the CPML",
                   "grid": [nx, ny, nz], "storage": "fp32", "kernel": r32["kernel"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if tr is None else round(tr["bytes"]),
                     "traffic_source": None if tr is None else tr["source"],
                     "kernel": r32["kernel"] + " " + r32["dominant"], "avg_launch_ms": dom["avg_ms"],
                     "algorithmic_bytes_per_launch": dom["algorithmic_bytes"], "cells_per_launch": dom["cells"],
                     "extra_bytes": dom["psi_bytes"], "extra": "CPML psi bytes per launch (not in achieved)",
                     "achieved_with_extra": round(dom["gbs_with_psi"], 1)},
        "effective_yee_gbs": round(r32["gcells"] * 72, 1),
        "effective_yee_frac": round(r32["gcells"] * 72 / HBM_PEAK_GBS, 4),
        "fp16_store": None if r16 is None else {
            "value": round(r16["gcells"], 3), "unit": "Gcells/s", "speedup_vs_fp32": round(r16["gcells"] / r32["gcells"], 3),
            "dominant_kernel_gbs": r16["groups"][r16["dominant"]]["gbs"],
            "purcell_error": None if pur is None else pur.get("fp16_zone_vs_fp64_error",
                                                              pur["fp16_zone_vs_fp32_error"]),
            "purcell_error_basis": None if pur is None else (
                ("FP64 oracle" if "fp16_zone_vs_fp64_error" in pur else "FP32 storage")
                + ", FP16 storage with the FP32 zone around the dipole"),
            "purcell_error_no_zone": None if pur is None else pur.get("fp16_vs_fp64_error",
                                                                      pur["fp16_vs_fp32_error"])},
        "purcell": pur,
        "other_schedule_fp32": None if ralt is None else {
            "kernel": ralt["kernel"], "value": round(ralt["gcells"], 3), "unit": "Gcells/s",
            "groups": ralt["groups"]},
        "adjoint_forward_ratio": None if adj is None else round(adj["ratio"], 3),
        "adjoint": adj,
        "cpu_baseline": cpu,
        "kernel_groups_fp32": r32["groups"],
    }
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        copy_calibration(nx * ny * nz * 6)


def copy_calibration(n):
    """Three device-to-device copies of n FP32 values: a kernel whose HBM
    traffic is known exactly (4n bytes read, 4n written per launch)."""
    import torch
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {3} launches, read {4 * n} B and write {4 * n} B per launch")


if __name__ == "__main__":
    main()
