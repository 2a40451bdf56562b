"""Builders of the workloads the stepping path is measured and checked on.

Each builder returns a `Workload`: the dtm_config, per-component material
volumes, probe lists and DFT frequencies.  `apply(w, sim)` installs the same
workload on a `Solver` or on anything with the same methods, so one
definition drives every consumer.  Sizes are parameters, so tests use
shrunken copies of the same physics.

Units follow the reference (c = 1, mu = 1, lengths in micrometres).
"Courant 0.5" means dt*sqrt(1/dx^2+1/dy^2+1/dz^2) = 0.5 (the (2,4) bound is
6/7, DTmxw.cu:332).
"""
from dataclasses import dataclass, field

import numpy as np

from . import abi

C_SI = 2.99792458e14  # um/s: converts rad/s to rad per (um/c)


@dataclass
class Workload:
    name: str
    cfg: abi.Config
    mats: tuple = (None, None, None)
    sensors: list = field(default_factory=list)   # (fld, x, y, z)
    probes: list = field(default_factory=list)    # (comp, i, j, k)
    omegas: np.ndarray = None
    design: tuple = None                           # (i0, j0, k0, eps array)
    steps: int = 0
    notes: dict = field(default_factory=dict)


def apply(w, sim):
    if any(m is not None for m in w.mats):
        sim.set_materials(*w.mats)
    if w.design is not None:
        i0, j0, k0, eps = w.design
        sim.set_design_eps(i0, j0, k0, eps)
    for s in w.sensors:
        sim.add_sensor(*s)
    for p in w.probes:
        sim.add_probe(*p)
    if w.omegas is not None and len(w.omegas):
        sim.set_dft(w.omegas)
    return sim


def dt_for(dx, dy, dz, courant=0.5):
    return courant / np.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2)


def component_positions(shape, comp, dx, dy, dz):
    """Physical (x, y, z) of component comp in every cell (lattice of
    include/dtmaxwell4.h: half-coordinate h -> h*d/2)."""
    off = [(0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)][comp]
    nx, ny, nz = shape
    x = (2 * np.arange(nx) + off[0]) * 0.5 * dx
    y = (2 * np.arange(ny) + off[1]) * 0.5 * dy
    z = (2 * np.arange(nz) + off[2]) * 0.5 * dz
    return x[:, None, None], y[None, :, None], z[None, None, :]


def sphere_mats(shape, d, centre, radius, mat_id, base=None):
    """Per-E-component material volumes of a sphere (component positions
    tested individually, as get_mat_indev is per component, defs.h:177-179)."""
    out = []
    for c in range(3):
        x, y, z = component_positions(shape, c, *d)
        r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
        m = np.zeros(shape, dtype=np.uint8) if base is None else base[c].copy()
        m[np.broadcast_to(r2 <= radius ** 2, shape)] = mat_id
        out.append(m)
    return tuple(out)


def seeded_points(n, shape, margin, avoid, avoid_r, seed=1234):
    rng = np.random.RandomState(seed)  # MT19937
    pts = []
    if any(s - 2 * margin <= 0 for s in shape):
        raise ValueError("no room for sensors: grid too small for the margin")
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 100000:
            raise ValueError("could not place sensors away from the source")
        p = [int(rng.randint(margin, s - margin)) for s in shape]
        if sum((p[a] - avoid[a]) ** 2 for a in range(3)) >= avoid_r ** 2:
            pts.append(tuple(p))
    return pts


# ------------------------------------------------------------------ workloads

def vacuum_dipole(n=128, npml=10, steps=1000, cells_per_lambda=20):
    """Plain FP32 vacuum box n^3, dx = lambda/20, Courant 0.5, CPML on all faces,
    soft Ez point source at the centre with a Gaussian-modulated sine
    (f0 = c/(20 dx), tau = 3/f0, t0 = 4 tau), Ez at 8 seeded sensors."""
    lam = 1.0
    dx = lam / cells_per_lambda
    dt = dt_for(dx, dx, dx)
    cfg = abi.default_config(n, n, n, dx, dx, dx, dt)
    cfg.npml[:] = [npml] * 6
    f0 = 1.0 / lam
    c = n // 2
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EZ, c, c, c
    so.waveform = abi.WAVE_GAUSS_SINE
    so.amp, so.w, so.tw, so.t0 = 1.0, 2 * np.pi * f0, 3.0 / f0, 12.0 / f0
    pts = seeded_points(8, (n, n, n), npml + 4, (c, c, c), 4)
    sensors = [("Ez",) + p for p in pts]
    return Workload("vacuum_dipole", cfg, sensors=sensors, steps=steps,
                    notes=dict(dx=dx, dt=dt, f0=f0, source=(c, c, c)))


def plane_wave_sphere(nx=512, ny=256, nz=256, npml=12, inset=16, radius=40, eps=4.0, steps=4000,
                      cells_per_lambda=20, storage=abi.STORE_FP32, periods_w=2.0, periods_t0=8.0):
    """TF/SF plane wave (x-propagating, Ez-polarised, Gaussian pulse) on a
    dielectric sphere (radius 40 cells, eps 4) in a 512x256x256 grid with a
    12-cell CPML and the TF/SF box 16 cells inside the CPML."""
    lam = 1.0
    dx = lam / cells_per_lambda
    dt = dt_for(dx, dx, dx)
    cfg = abi.default_config(nx, ny, nz, dx, dx, dx, dt)
    cfg.npml[:] = [npml] * 6
    cfg.storage = storage
    cfg.nmat = 2
    cfg.mat[1].eps_inf = eps
    b = npml + inset
    t = cfg.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_PLANE, abi.WAVE_GAUSS_SINE, 0
    t.box[:] = [b * dx, (nx - b) * dx, b * dx, (ny - b) * dx, b * dx, (nz - b) * dx]
    f0 = 1.0 / lam
    t.w = 2 * np.pi * f0
    t.cSrc, t.Imp, t.amp = 1.0, 1.0, 1.0
    t.theta, t.phi, t.psi = np.pi / 2, 0.0, -np.pi / 2   # k = +x, E = +z
    t.tw, t.t0 = periods_w / f0, periods_t0 / f0
    t.tStop = 1e30
    t.pos[:] = [b * dx, 0.0, 0.0]
    centre = (0.5 * nx * dx, 0.5 * ny * dx, 0.5 * nz * dx)
    mats = sphere_mats((nx, ny, nz), (dx, dx, dx), centre, radius * dx, 1)
    if storage == abi.STORE_FP16:
        # the sphere focuses the pulse to |H| = 16 behind it (step ~1700 at
        # the full size): 2^10 leaves the half range (65504 / 2^10 = 64) 4x
        # above that
        cfg.scale_e_log2[:] = [10, 10, 10]
        cfg.scale_h_log2[:] = [10, 10, 10]
    c = (nx // 2, ny // 2, nz // 2)
    sensors = [("Ez", c[0] - radius - 4, c[1], c[2]), ("Ez", c[0] + radius + 4, c[1], c[2]),
               ("Ez", c[0], c[1], c[2])]
    return Workload("plane_wave_sphere", cfg, mats=mats, sensors=sensors, steps=steps,
                    notes=dict(dx=dx, dt=dt, f0=f0))


def drude_lorentz_purcell(n=384, npml=12, radius=40, gap=4, nfreq=64, steps=4000, lam0_um=0.6,
                          cells_per_lambda=30, storage=abi.STORE_FP32, box_half=2):
    """Dispersive Purcell run: silver-like Drude pole (eps_inf=3.7,
    wp=1.37e16 rad/s, gamma=2.7e13 rad/s) plus one Lorentz pole on a sphere,
    Ez dipole `gap` cells from the surface, FP64 DFT bins on 0.5-1.5 f0.
    The Lorentz pole (d_eps=0.5, w_L=2pi c/0.25um, gamma_L=0.1 w_L) is this
    build's choice; the dipole is the reference's DipoleTFSF (scattered field
    inside a small box, signal.hpp:51-97)."""
    dx = lam0_um / cells_per_lambda
    dt = dt_for(dx, dx, dx)
    cfg = abi.default_config(n, n, n, dx, dx, dx, dt)
    cfg.npml[:] = [npml] * 6
    cfg.storage = storage
    cfg.nmat = 2
    m = cfg.mat[1]
    m.eps_inf = 3.7
    m.npoles = 2
    wp = 1.37e16 / C_SI
    gam = 2.7e13 / C_SI
    m.pole[0].setDrude(1.0, wp, gam / 2, 0.0)            # b1 = gamma (LorenzDisp::setDrude)
    wl = 2 * np.pi / 0.25
    m.pole[1].set(0.5, wl, 0.05 * wl, 0.0)               # b1 = 2*gp = 0.1 w_L
    centre = np.array([0.5 * n * dx, 0.5 * n * dx, 0.5 * n * dx])
    mats = sphere_mats((n, n, n), (dx, dx, dx), centre, radius * dx, 1)
    # dipole on the +z axis, `gap` cells above the surface, at an Ez node
    ci = n // 2
    kd = int(round(centre[2] / dx + radius + gap))
    pos = np.array([(ci + 0.5) * dx, (ci + 0.5) * dx, kd * dx])  # Ez(i,j,k) at ((i+.5)dx,(j+.5)dx,k dx)
    t = cfg.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_DIPOLE, abi.WAVE_GAUSS_SINE, 1
    h = box_half * dx
    t.box[:] = [pos[0] - h, pos[0] + h, pos[1] - h, pos[1] + h, pos[2] - h, pos[2] + h]
    w0 = 2 * np.pi / lam0_um
    # dipole moment scaled so the near field at the box faces is O(1); a
    # Gaussian width of half a period puts spectral weight > 0.5 of the peak
    # across the 0.5-1.5 f0 band the Purcell spectrum covers
    t.w, t.cSrc, t.Imp, t.amp = w0, 1.0, 1.0, (box_half * dx) ** 3
    t.tw, t.t0 = 0.5 * 2 * np.pi / w0, 4 * 0.5 * 2 * np.pi / w0
    t.tStop = 1e30
    if storage == abi.STORE_FP16:
        cfg.scale_e_log2[:] = [10, 10, 10]
        cfg.scale_h_log2[:] = [10, 10, 10]
    om = w0 * np.linspace(0.5, 1.5, nfreq)
    probes = [(abi.EZ, ci, ci, kd)]
    return Workload("drude_lorentz_purcell", cfg, mats=mats, probes=probes, omegas=om, steps=steps,
                    notes=dict(dx=dx, dt=dt, w0=w0, dipole=(ci, ci, kd), pos=pos, box_half=box_half))


def dipole_zone(w, margin=8):
    """FP32 zone for FP16 storage around a dipole TF/SF box: the cells within
    box_half + margin of the dipole cell (clipped to the grid), where the
    total field is large; as (i0, i1, j0, j1, k0, k1)."""
    c = w.notes["dipole"]
    h = int(w.notes["box_half"]) + margin
    n = (w.cfg.nx, w.cfg.ny, w.cfg.nz)
    out = []
    for a in range(3):
        out += [max(c[a] - h, 0), min(c[a] + h + 1, n[a])]
    return tuple(out)


def gyro_slab(nx=256, ny=256, nz=512, npml=12, thick=64, steps=8000, lam0_um=1.5, cells_per_lambda=40,
              storage=abi.STORE_FP32, wp_rad_s=2e15, nu_rad_s=1e13, wc_ratio=0.2, eps_inf=4.0):
    """Gyrotropic magneto-optical slab: slab `thick` cells along z with
    eps_inf=4 and a magnetised-plasma ADE (wp=2e15 rad/s, nu=1e13 rad/s, bias
    along z with wc=0.2 wp), normally incident TF/SF Gaussian pulse along z
    (x-polarised), transmitted field probes for the Faraday rotation."""
    dx = lam0_um / cells_per_lambda
    dt = dt_for(dx, dx, dx)
    cfg = abi.default_config(nx, ny, nz, dx, dx, dx, dt)
    cfg.npml[:] = [0, 0, 0, 0, npml, npml]
    cfg.bc[0] = cfg.bc[1] = abi.BC_PERIODIC
    cfg.storage = storage
    cfg.nmat = 2
    m = cfg.mat[1]
    m.eps_inf = eps_inf
    m.gyro = 1
    m.gyro_wp = wp_rad_s / C_SI
    m.gyro_nu = nu_rad_s / C_SI
    m.gyro_wc[:] = [0.0, 0.0, wc_ratio * wp_rad_s / C_SI]
    z0 = nz // 2 - thick // 2
    mats = []
    for c in range(3):
        mm = np.zeros((nx, ny, nz), dtype=np.uint8)
        mm[:, :, z0:z0 + thick] = 1
        mats.append(mm)
    b = npml + 8
    t = cfg.tfsf
    t.kind, t.waveform, t.sf_inside = abi.TFSF_PLANE, abi.WAVE_GAUSS_SINE, 0
    # TF region: a slab in z spanning the whole periodic x/y extent
    t.box[:] = [-1e9, 1e9, -1e9, 1e9, b * dx, (nz - b) * dx]
    w0 = 2 * np.pi / lam0_um
    t.w, t.cSrc, t.Imp, t.amp = w0, 1.0, 1.0, 1.0
    t.theta, t.phi, t.psi = 0.0, np.pi / 2, 0.0      # k = +z; E along -phi_hat... = x
    t.tw, t.t0 = 3 * 2 * np.pi / w0, 4 * 3 * 2 * np.pi / w0
    t.tStop = 1e30
    t.pos[:] = [0.0, 0.0, b * dx]
    ci, cj = nx // 2, ny // 2
    kt = z0 + thick + (nz - b - (z0 + thick)) // 2
    probes = [(abi.EX, ci, cj, kt), (abi.EY, ci, cj, kt)]
    om = np.array([w0])
    return Workload("gyro_slab", cfg, mats=tuple(mats), probes=probes, omegas=om, steps=steps,
                    notes=dict(dx=dx, dt=dt, w0=w0, z0=z0, thick=thick))


def inverse_design(nx=400, ny=200, nz=200, npml=12, design=(64, 64, 32), steps=6000, seed=7,
                   cells_per_lambda=20, monitor=16):
    """Adjoint inverse design: design region of density rho~U[0,1] (seeded,
    3-cell box-smoothed), eps=1+11 rho; soft Ez dipole; objective
    sum |DFT Ez|^2 over a monitor plane at one frequency."""
    lam = 1.0
    dx = lam / cells_per_lambda
    dt = dt_for(dx, dx, dx)
    cfg = abi.default_config(nx, ny, nz, dx, dx, dx, dt)
    cfg.npml[:] = [npml] * 6
    f0 = 1.0 / lam
    dnx, dny, dnz = design
    i0, j0, k0 = nx // 2 - dnx // 2, ny // 2 - dny // 2, nz // 2 - dnz // 2
    rng = np.random.RandomState(seed)
    rho = rng.uniform(0, 1, size=design)
    # 3-cell box smoothing
    from scipy.ndimage import uniform_filter
    rho = uniform_filter(rho, size=3, mode="nearest")
    eps = 1.0 + 11.0 * rho
    src = (i0 - max(4, dnx // 8), ny // 2, nz // 2)
    cfg.nsoft = 1
    so = cfg.soft[0]
    so.comp, so.i, so.j, so.k = abi.EZ, src[0], src[1], src[2]
    so.waveform = abi.WAVE_GAUSS_SINE
    so.amp, so.w, so.tw, so.t0 = 1.0, 2 * np.pi * f0, 3.0 / f0, 12.0 / f0
    im = i0 + dnx + max(4, dnx // 8)
    half = monitor // 2
    probes = [(abi.EZ, im, ny // 2 - half + a, nz // 2 - half + b) for a in range(monitor) for b in range(monitor)]
    return Workload("inverse_design", cfg, probes=probes, omegas=np.array([2 * np.pi * f0]),
                    design=(i0, j0, k0, eps), steps=steps,
                    notes=dict(dx=dx, dt=dt, f0=f0, rho=rho, design_origin=(i0, j0, k0)))
