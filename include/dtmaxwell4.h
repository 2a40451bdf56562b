/*
 * dtmaxwell4.h — C ABI of the MI355X FDTD stepping path.
 *
 * Everything here is plain C: integers, doubles, floats and pointers to host
 * memory.  No HIP or torch type crosses this boundary.  Two surfaces are
 * exported by libdtmaxwell4.so:
 *
 *  (1) the legacy DTmxw surface that the reference's SWIG module binds
 *      (mxw.i:1-36, run.py:11-40): _main, calcStep, the globals Tsteps and
 *      shotpoint, TFSFsrc_set/TFSFsrc_check and the grid constants;
 *  (2) an explicit handle API (dtm_*) with integer error codes instead of
 *      exit(): create / step / read sensors, DFT bins and fields / adjoint /
 *      destroy.
 *
 * Units follow the reference: c0 = 1, mu = 1, eps relative
 * (py_consts.h:71, defs.h:150-152).  Field arrays on the boundary are
 * C-ordered [nx][ny][nz] (z fastest).
 *
 * Lattice: the reference places fields on half-cell integer coordinates
 * (genAlg.py:22-24): E ("V") has exactly two odd coordinates, H ("T") exactly
 * one.  Cell (i,j,k) of this library holds
 *   Ex (2i, 2j+1, 2k+1)  Ey (2i+1, 2j, 2k+1)  Ez (2i+1, 2j+1, 2k)
 *   Hx (2i+1, 2j, 2k)    Hy (2i, 2j+1, 2k)    Hz (2i, 2j, 2k+1)
 * so the physical position of half-coordinate h along axis a is h*d_a/2,
 * as in DipoleTFSF (signal.hpp:59-61).
 */
#ifndef DTMAXWELL4_H
#define DTMAXWELL4_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ enums */
enum { DTM_EX = 0, DTM_EY, DTM_EZ, DTM_HX, DTM_HY, DTM_HZ };
enum { DTM_BC_ZERO = 0, DTM_BC_PERIODIC = 1 };
enum { DTM_STORE_FP32 = 0, DTM_STORE_FP16 = 1 };
enum { DTM_TFSF_NONE = 0, DTM_TFSF_DIPOLE = 1, DTM_TFSF_PLANE = 2 };
/* Source time functions.  RAISED_COS is the reference's window
 * 0.5(1-cos(kEnv*phase))*sin(phase) on 0<=kEnv*phase<=2pi (signal.hpp:35,71-77);
 * GAUSS_SINE is exp(-((tau-t0)/tw)^2) sin(w tau) with tau=phase/w
 * (the envelope of signal.hpp:55-57).                                     */
/* GAUSS is the reference's envelope pulse without a carrier,
 * exp(-((tau-t0)/tw)^2) (EnvelopeTime/EnvelopeDiv1/EnvelopeDiv2,
 * signal.hpp:55-57, used as the dipole moment in DipoleTFSF's alternative
 * branch, signal.hpp:84-87): the broadband pulse of the reference's fdip.dat
 * (its amplitude spectrum falls as exp(-w^2 tw^2/4) with tw = 0.1). */
enum { DTM_WAVE_RAISED_COS = 0, DTM_WAVE_GAUSS_SINE = 1, DTM_WAVE_GAUSS = 2 };
/* Absorbing layer: CPML (psi only in the slabs, kappa = 1, alpha = 0), or the
 * reference's split-field layer (genAlg.py:117-148): in the slabs each
 * component is the sum of one part per derivative axis, each part advanced
 * as part = k1 part +- k2 dif with k2 = 1/(1 + sigma dt/2), k1 = 2 k2 - 1
 * (DTmxw.cu:250-265).  Both use the same sigma grading. */
enum { DTM_PML_CPML = 0, DTM_PML_SPLIT = 1 };
/* sigma grading across a layer of p cells (2p half-cells), u half-cells from
 * its outer face:
 *  POLY       sigma_max ((2p - u) / 2p)^m, sigma_max = V ln(1/R) (m+1) / (2 p d);
 *  REFERENCE  PMLgamma_func (DTmxw.cu:219-228) with the index map of
 *             setPMLcoeffs for x (DTmxw.cu:251-253): sigma_max over the outer
 *             3 half-cells, sigma_max (1 - (u-3)/N)^m over N = 2p - 6
 *             half-cells, zero over the inner 3; sigma_max = V ln(1/R) (m+1)
 *             / (N d).  Needs p >= 4 on every face with a layer. */
enum { DTM_PML_PROFILE_POLY = 0, DTM_PML_PROFILE_REFERENCE = 1 };
enum {
  DTM_OK = 0,
  DTM_ERR_ARG = -1,      /* invalid argument / configuration */
  DTM_ERR_COURANT = -2,  /* dt*sqrt(sum 1/d^2) > 6/7 (DTmxw.cu:332) */
  DTM_ERR_DEVICE = -3,   /* HIP runtime error */
  DTM_ERR_NOMEM = -4,
  DTM_ERR_STATE = -5,    /* call not valid in the current state */
  DTM_ERR_LIMIT = -6     /* a fixed table size was exceeded */
};

#define DTM_MAX_MATS 64
#define DTM_MAX_POLES 4
#define DTM_MAX_SOFT 16

/* ------------------------------------------------------- reference types */

/* One Lorentz/Drude pole, chi(w) = (a0 - i w a1)/(b0 - i w b1 - b2 w^2).
 * Same fields and meaning as LorenzDisp (dispersion.cpp:5-22). */
typedef struct LorenzDisp {
  double a0, a1, b0, b1, b2;
} LorenzDisp;
void LorenzDisp_set(LorenzDisp* p, double Deps, double wp, double gp, double gp_);
void LorenzDisp_setDrude(LorenzDisp* p, double Deps, double wp, double gp, double gp_);

/* The reference's source parameter block, field for field (py_consts.h:73-91)
 * with ftype=float.  Exported as the global `shotpoint`. */
typedef struct TFSFsrc {
  float tStop;
  float srcXs, srcXa, srcXv, sphR;
  float BoxMs, BoxPs;
  float BoxMa, BoxPa;
  float BoxMv, BoxPv;
  float V_max;
  int start;
  float wavelength, w, k, Omega, kEnv, theta, phi;
  float cSrc;
  float Imp;
  float t0, tw, T;
} TFSFsrc;

/* Per-material coefficient block of the reference (defs.h:88-118).  Only
 * set_eps's algebra is used (gyrotropic coefficients); the kernels read the
 * library's own compact tables. */
typedef struct CoffStruct {
  float deps, depsXX, depsYY, depsZZ;
  float coffAnisoXY, coffAnisoXZ, coffAnisoYX, coffAnisoYZ, coffAnisoZX, coffAnisoZY;
  float gyrationXX, gyrationXY, gyrationXZ;
  float gyrationYX, gyrationYY, gyrationYZ;
  float gyrationZX, gyrationZY, gyrationZZ;
  int isDisp;
} CoffStruct;
/* CoffStruct::set_eps (defs.h:96-117): g <- g*w*dt, then the Cayley pair
 * deps = 2(2eps+G)^-1 and gyration = (eps+G/2)^-1 (eps-G/2).             */
void CoffStruct_set_eps(CoffStruct* c, double eps, int disp, double gx, double gy,
                        double gz, double w, double dt);

/* ----------------------------------------------------- library config */

/* Material: eps_inf, conductivity, up to DTM_MAX_POLES Lorentz/Drude poles
 * (the DispStruct of dispersion.cpp:41-82) and an optional magnetised-plasma
 * (gyrotropic) current: dJ/dt + nu J = wp^2 E + wc x J.                   */
typedef struct dtm_material {
  double eps_inf;
  double sigma;
  int npoles;
  LorenzDisp pole[DTM_MAX_POLES];
  int gyro;
  double gyro_wp, gyro_nu, gyro_wc[3];
} dtm_material;

/* TF/SF source.  `sf_inside`=1 reproduces the reference: the box interior is
 * the scattered-field region (signal.hpp:153-160, genAlg.py:219-221).  A
 * plane wave normally uses sf_inside=0 (total field inside).  Incident fields
 * are evaluated in double precision. */
typedef struct dtm_tfsf {
  int kind;        /* DTM_TFSF_* */
  int waveform;    /* DTM_WAVE_* */
  int sf_inside;
  double box[6];   /* BoxMs, BoxPs, BoxMa, BoxPa, BoxMv, BoxPv (physical) */
  double w;        /* angular frequency */
  double kEnv;     /* raised-cosine envelope ratio Omega/w */
  double cSrc;     /* phase velocity used in the source */
  double Imp;      /* impedance E/H of the incident wave */
  double theta, phi, psi; /* plane wave: propagation angles, polarisation;
                            dipole: axis (sin th cos ph, sin th sin ph, cos th),
                            th = 0 the reference's z dipole (signal.hpp:70-72) */
  double t0, tw;   /* Gaussian envelope centre and width */
  double amp;      /* amplitude (the reference's A=1) */
  double tStop;    /* corrections are applied while the gate allows it */
  double pos[3];   /* plane wave phase origin (x,y,z); dipole uses box centre */
} dtm_tfsf;

/* Soft current source J_c(t) = amp*wave(t) on one E component of one cell:
 * E^{n+1} -= (dt/eps) J(t_{n+1/2}). */
typedef struct dtm_soft_source {
  int comp;        /* DTM_EX..DTM_EZ */
  int i, j, k;
  int waveform;    /* DTM_WAVE_* */
  double amp, w, t0, tw, kEnv;
} dtm_soft_source;

/* Time-step schedule.  TWOPASS: one sweep for H, one for E (72 B per
 * cell-update in FP32).  FUSED: H and E of one step in a single x-marching
 * sweep over double-buffered fields (48 B per cell-update plus halo
 * recompute); needs a second copy of the fields. */
enum { DTM_SWEEP_TWOPASS = 0, DTM_SWEEP_FUSED = 1 };

typedef struct dtm_config {
  int nx, ny, nz;
  double dx, dy, dz, dt;
  int bc[3];              /* DTM_BC_* per axis */
  int npml[6];            /* CPML thickness in cells: x-, x+, y-, y+, z-, z+ */
  double pml_order;       /* grading exponent m (reference: 4, DTmxw.cu:224) */
  double pml_R;           /* target reflection (reference: 1e-4, DTmxw.cu:225) */
  double pml_vmax;        /* V_max (reference: shotpoint.V_max = c) */
  int pml_kind;           /* DTM_PML_CPML or DTM_PML_SPLIT */
  int storage;            /* DTM_STORE_FP32 / DTM_STORE_FP16 */
  int scale_e_log2[3];    /* FP16 storage: E_c stored as E_c * 2^s */
  int scale_h_log2[3];
  int nmat;               /* materials referenced by the index volumes */
  dtm_material mat[DTM_MAX_MATS];
  dtm_tfsf tfsf;
  int ntime;              /* steps per calcStep (reference Ntime): TF/SF gate
                             granularity and legacy sensor-write cadence   */
  int nsoft;
  dtm_soft_source soft[DTM_MAX_SOFT];
  int sweep;              /* DTM_SWEEP_*: time-step schedule (same results) */
  int pml_profile;        /* DTM_PML_PROFILE_*: sigma grading across the layer */
} dtm_config;

/* Fill a config with the reference's defaults (dx=dy=dz=0.005, dt=0.002,
 * m=4, R=1e-4, V_max=1, vacuum material 0, no source). */
void dtm_config_defaults(dtm_config* cfg, int nx, int ny, int nz);

/* ---------------------------------------------------- explicit handle API */
typedef struct dtm_solver dtm_solver;

/* Last error message of the calling thread ("" if none). */
const char* dtm_last_error(void);
/* Library version string. */
const char* dtm_version(void);
/* Number of HIP devices visible (0 when none). */
int dtm_device_count(void);

/* Validates the config (Courant <= 6/7 for the (2,4) stencil, sizes, tables)
 * and allocates device state on the current HIP device.  Fields start at 0. */
int dtm_create(const dtm_config* cfg, dtm_solver** out);
void dtm_destroy(dtm_solver* s);

/* Per-component material index volumes, each nx*ny*nz bytes, C order.  Any
 * pointer may be NULL (material 0).  Replaces the texture/get_mat_indev
 * lookup of defs.h:128-182 with a compact per-cell class byte. */
int dtm_set_materials(dtm_solver* s, const uint8_t* mat_ex, const uint8_t* mat_ey,
                      const uint8_t* mat_ez);
/* Continuous permittivity override on a box (inverse design): eps[c] for the
 * cells [i0,i0+ni)x[j0,j0+nj)x[k0,k0+nk), array of ni*nj*nk doubles applied
 * to all three E components of a cell.  Non-dispersive cells only. */
int dtm_set_design_eps(dtm_solver* s, int i0, int j0, int k0, int ni, int nj, int nk,
                       const double* eps);

/* Point sensors (sensor.h:6-84): `fld` in "Ex".."Hz"; (x,y,z) is the
 * reference's Yee-node coordinate; the value is the reference's 4-point (E)
 * or 2-point (H) average.  Recorded on the device every step. */
int dtm_add_sensor(dtm_solver* s, const char* fld, int x, int y, int z);
/* Raw single-component probes: comp DTM_EX..DTM_HZ at cell (i,j,k). */
int dtm_add_probe(dtm_solver* s, int comp, int i, int j, int k);
/* Running FP64 DFT of every probe: F(w) = sum_n f(t_n) exp(i w t_n) dt. */
int dtm_set_dft(dtm_solver* s, int nfreq, const double* omegas);

/* Advance n full steps (H half-step then E half-step, genAlg.py:425-499). */
int dtm_step(dtm_solver* s, int n);
/* Steps done so far. */
long long dtm_time_step(const dtm_solver* s);
/* Block until queued work completes. */
int dtm_sync(dtm_solver* s);

/* Probe/sensor series: out[step][probe], nsteps_max rows; returns rows. */
int dtm_probe_count(const dtm_solver* s);
int dtm_probe_read(dtm_solver* s, float* out, int nsteps_max);
/* Newest row of the probe series (the values after the last step), one
 * float per probe: what Sensor::write reads once per calcStep
 * (sensor.h:51-83, diamond.cu:473-475).  Returns 1, or 0 before the first
 * step or without probes. */
int dtm_probe_last(dtm_solver* s, float* out);
/* DFT bins: out[probe][freq][re,im] doubles. */
int dtm_dft_read(dtm_solver* s, double* out);

/* Field access: comp DTM_EX..DTM_HZ, buf holds nx*ny*nz floats. */
int dtm_field_read(dtm_solver* s, int comp, float* buf);
int dtm_field_write(dtm_solver* s, int comp, const float* buf);
/* One component written as an aivlib .arr file, the snapshot format of the
 * reference's SeismoDrops::drop (drop.cu:84-107): int 12, three zero ints,
 * int 3 (dimensions), int 4 (bytes per value), the extents nz, ny, nx, then
 * nx*ny*nz floats with z fastest. */
int dtm_write_arr(dtm_solver* s, int comp, const char* path);
/* Save / resume: the whole time-stepped state (current fields, CPML psi,
 * split-field parts, dispersive and plasma currents), the step count, the
 * FP64 DFT bins and flux totals, to a file; dtm_load_state restores it into
 * a solver created with the same config, materials, probes, DFT and flux
 * boxes, which then continues bit for bit as the saved one would have.  The
 * probe and flux series recorded before the save are not in the file.
 * Replaces the reference's --swap_dir window swap files (window.hpp:177-267)
 * as the restart path; no file format of the reference is kept. */
int dtm_save_state(dtm_solver* s, const char* path);
int dtm_load_state(dtm_solver* s, const char* path);
/* Field energy 1/2 sum(eps_inf E^2 + H^2) dV of the stored state (E^n and
 * H^{n-1/2}), summed in FP64 on the device in a fixed order (the same value
 * on every call).  dtm_energy: the whole grid; dtm_energy_box: the cells
 * [i0,i1) x [j0,j1) x [k0,k1).  The design box's eps is used where it is set.
 * No counterpart in the reference. */
int dtm_energy(dtm_solver* s, double* out);
int dtm_energy_box(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1, double* out);
/* dtm_energy_box for callers that cannot allocate device memory: the same
 * kernels and the same value, with the per-plane partials in a static device
 * array of the library (8192 doubles plus the result, held by every device
 * the library runs on until the process ends) instead of memory the library
 * allocates on first use.  Calls no hipMalloc or hipFree; boxes spanning
 * more than 8192 x planes get DTM_ERR_LIMIT.  Every call on a device shares
 * that device's array: calls that overlap on one device race.  No
 * counterpart in the reference; for callers that cannot allocate (code
 * under stream capture, tight serving loops). */
int dtm_energy_box_noalloc(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1, double* out);

/* Where a plane-wave TF/SF source along x is applied: on (the default)
 * inside the H and E sweeps, or inside the fused step, added to the
 * derivatives the face terms of genAlg.py:219-221 belong to, from per-plane
 * factors; off: in a separate correction kernel after each sweep (before and
 * after the fused step).  Both give the same fields up to rounding (the
 * updates are linear in those terms).  Other plane waves, dipoles and a
 * plane wave with the FP32 zone always use the separate kernel. */
int dtm_set_tfsf_fold(dtm_solver* s, int on);

/* ------------------------------------------------------------ FP32 zone */
/* FP16 storage with a box of cells [i0,i1) x [j0,j1) x [k0,k1) whose six
 * components are also kept in FP32: every kernel that reads or writes a
 * field value of a zone cell uses the FP32 copy (the half array holds its
 * rounded value, so readers that do not know the zone, such as the flux and
 * energy sums, see FP16 precision there).  Put it around a dipole TF/SF box
 * and its probes: the total field there is large and its FP16 rounding
 * otherwise dominates a Purcell factor's error.  FP16 storage and the
 * two-pass sweeps only, set before the first step, once, at most 2^27 cells.
 * dtm_field_read / dtm_field_write and the state files include the zone.
 * No counterpart in the reference (its fields are FP32 ftype, params.h:26-29). */
int dtm_set_fp32_zone(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1);

/* ------------------------------------------------------- Poynting flux */
/* FP64 Poynting-flux sums through the six faces of a box of cells
 * [i0,i1) x [j0,j1) x [k0,k1), recorded on the device every step: the
 * outward flux of S = E x H through the faces x-, x+, y-, y+, z-, z+ at
 * t = (n+1/2) dt, from H^{n+1/2} and (E^n + E^{n+1})/2 averaged onto the face
 * planes (second order in space and time).  Every face plane must lie on
 * planes 1 .. n-1 of its axis.  Up to 8 boxes, added before the first step;
 * returns the box id.  No counterpart in the reference (it writes point
 * sensors only, sensor.h). */
int dtm_add_flux_box(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1);
/* Flux series: out[step][box*6 + face] doubles, nsteps_max rows; returns rows. */
int dtm_flux_read(dtm_solver* s, double* out, int nsteps_max);
/* Running FP64 time integrals sum_n flux dt: out[box*6 + face]. */
int dtm_flux_total(dtm_solver* s, double* out);

/* -------------------------------------------------------------- adjoint */
/* Gradient of J = sum_m |F_m|^2, F_m = sum_{n<nsteps} Ez_m^{n+1} e^{i omega (n+1) dt} dt,
 * with Ez_m the raw Ez of monitor cell m (mon_ijk: nmon (i,j,k) triples),
 * with respect to the design box's inverse permittivity (dtm_set_design_eps).
 * Runs nsteps forward from step 0 recording the design-box E, then the
 * discrete adjoint backward (the transposed leapfrog step, which is the
 * forward step on (eps^-1 lambda_E, -lambda_H)).  J_out gets J;
 * grad_inv_eps gets dJ/d(1/eps) per design cell (C order, ni*nj*nk doubles).
 * Afterwards the solver is back at step 0 with zero fields.  Needs FP32
 * storage and no dispersive or gyrotropic cells; holds (nsteps+1)*3*ni*nj*nk
 * floats of device memory for the call.  No counterpart in the reference
 * (it has no adjoint); the explicit API's dtm_adjoint entry. */
int dtm_adjoint_design_gradient(dtm_solver* s, int nsteps, double omega, int nmon, const int* mon_ijk,
                                double* J_out, double* grad_inv_eps);
/* Memory scheme of dtm_adjoint_design_gradient.  every = 0 (default) keeps
 * the design-box E of every step.  every = K > 0 keeps the whole time-stepped
 * state every K steps instead and, segment by segment on the way back,
 * recomputes each K steps forward to rebuild their design-box E.  The last
 * L = 1 + (state bytes / K steps of snapshots) segments keep their snapshots
 * from the forward pass instead of a checkpoint: device memory at most
 * (ceil(nsteps/K) + 1) states plus K+1 design-box snapshots, for one more
 * forward pass less L segments.  For grids whose design box is a large part
 * of the domain.
 * No counterpart in the reference (it has no adjoint). */
int dtm_adjoint_set_checkpoints(dtm_solver* s, int every);

/* ---------------------------------------------------------------- timing */
/* Time `nsteps` steps with HIP events on the solver's stream; returns ms. */
int dtm_time_steps(dtm_solver* s, int nsteps, double* ms);
/* Name of the kernel variant used for the E/H sweep. */
const char* dtm_kernel_name(const dtm_solver* s);
/* Per-launch HIP-event timing of the sweep kernels, recorded on the solver's
 * stream while enabled.  Groups: 0 H sweep (fused schedule: the whole step),
 * 1 unused, 2 E sweep, 3 unused, 4 TF/SF, ADE, sources and probes.
 * dtm_kernel_times returns the summed ms and launch counts per group and
 * resets them. */
int dtm_set_kernel_timing(dtm_solver* s, int on);
int dtm_kernel_times(dtm_solver* s, double* ms, long long* count);
/* Cells covered by group g's launches. */
long long dtm_group_cells(const dtm_solver* s, int g);

/* ------------------------------------------------------- legacy surface */
/* Globals of the reference (DTmxw.cu:729-730, signal.hpp:12, py_consts.h). */
extern int Tsteps;
extern TFSFsrc shotpoint;
/* Field snapshots after every calcStep, the reference's compile-time
 * DROP_DATA build (py_consts.h:19, diamond.cu:469-470, drop.cu:75-107) as a
 * run-time switch: bit c (0..5: Ex..Hz) writes <drop_dir>/rag<comp>_<n>.arr
 * (n: the calcStep count, %05d) in the aivlib layout of dtm_write_arr.  0 (the
 * default) is the reference's default build; 4 (Ez) is its DROP_DATA build
 * (drop.cu:43-44).  _main removes such files left in drop_dir, as the
 * reference's SeismoDrops::init does. */
extern int DropData;
extern int GridNx, GridNy, GridNz;     /* true x, y, z extents (cells) */
extern float ds, da, dv, dt;           /* py_consts.h:71 */
extern const int NDT;                  /* py_consts.h:1 */
extern int Npmlx, Npmly, Npmlz;        /* CPML thickness per side in cells */
extern int Ntime;                      /* steps per calcStep */
extern int NyBloch, Nzw, Np;

/* TFSFsrc::set / ::check (signal.hpp:99-133). */
void TFSFsrc_set(TFSFsrc* s, double LightSpeed, double impedance, double lmb);
void TFSFsrc_check(TFSFsrc* s);
/* int _main(int argc, char** argv) (DTmxw.cu:731-834), headless part. */
int _main(int argc, char** argv);
/* int calcStep() (diamond.cu:335-506): advance Ntime steps, write sensors,
 * print the rate line. */
int calcStep(void);

#ifdef __cplusplus
}
#endif
#endif /* DTMAXWELL4_H */
