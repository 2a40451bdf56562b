/*
 * fdtd_oracle.c — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, CPU restatement of the reference's FDTD stepping path, used as
 * the checker for the HIP library.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Built twice by oracle/Makefile:
 *   liboracle64.so  — ORC_REAL=double: the FP64 oracle;
 *   liboracle32.so  — ORC_REAL=float : the same equations in FP32 with OpenMP,
 *                     the "plain multithreaded CPU run" reported as the CPU
 *                     baseline.
 *
 * What each part restates (reference file:line) and what pins it
 * (tests/test_oracle_*.py):
 *   - (2,4) staggered E/H update, genAlg.py:182-301 (stencil :196-198,
 *     signs :246-247, plain update :269-272), constants defs.h:3-13.
 *     Pinned by the closed-form (2,4) numerical dispersion relation of a
 *     periodic eigenmode, the leapfrog energy invariant, div E = 0, and the
 *     6/7 stability bound (DTmxw.cu:332).
 *   - boundaries: zero padding outside the grid (genAlg.py:38 "zerov",
 *     defs.h:40 isOutS) or periodic wrap (defs.h:204).  Pinned by the
 *     periodic eigenmode (wrap) and the leapfrog invariant in a zero-padded
 *     box (conservative only if the padding is the exact transpose pair).
 *   - absorbing layer: CPML with the reference's sigma grading
 *     (DTmxw.cu:220-249: sigma_max = V_max ln(1/R) (m+1)/(2L), (1-i/N)^m).
 *     Pinned by the sigma closed form and a reflection bound vs a large box.
 *     The reference profile (PMLgamma_func as setPMLcoeffs indexes it along
 *     x, DTmxw.cu:219-253) is pinned by its flat outer and zero inner
 *     3-half-cell margins, the round trip of its graded part and the same
 *     reflection bound, in the CPML and the split-field layer.
 *   - TF/SF corrections, genAlg.py:207-234 with inSF signal.hpp:153-160;
 *     dipole incident field signal.hpp:51-97, plane wave :14-50.  Pinned by
 *     zero leakage outside the box without a scatterer and the closed-form
 *     dipole field identities (far field E.n = 0, |E| = c|B|).
 *   - Lorentz/Drude ADE, dispersion.cpp:41-99.  Pinned by the discrete
 *     eps(w) of the recursion (Z-transform) converging at second order to
 *     the continuous Lorentz/Drude permittivity.
 *   - sensors, sensor.h:51-83.  Pinned by the interpolation formula on an
 *     analytic field.
 *   - FP64 DFT (new on the device; the reference post-processes with an
 *     external script, calc_purcell.sh:2).  Pinned by the DFT of an
 *     analytic waveform.
 *   - magnetised-plasma ADE (no reference counterpart: Gyration() is
 *     disabled, genAlg.py:303-304).  Pinned by second-order convergence of
 *     the current update to dJ/dt + nu J = wp^2 E + wc x J and by the
 *     losslessness of its rotation part (tests/test_oracle_media.py).
 *   - CoffStruct::set_eps (defs.h:96-117), kept reference-literal including
 *     its coffAnisoXZ term; pinned by the Cayley identities where gx = 0 and
 *     by the closed form of the XZ deviation.
 *   - Purcell post-processing (purcell.gp:2): image-dipole closed form and
 *     the rows of the reference's own spectra (tests/golden).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/dtmaxwell4.h"

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define MAXPROBES 4096

typedef struct {
  int kind;  /* 0 raw component, 1 reference sensor (sensor.h) */
  int comp;  /* DTM_EX..DTM_HZ */
  int i, j, k;
} orc_probe;

typedef struct orc_sim {
  dtm_config cfg;
  int nx, ny, nz;
  size_t n;
  real* E[3];
  real* H[3];
  real* psiE[3][3]; /* [component][derivative axis] (NULL on own axis) */
  real* psiH[3][3];
  /* CPML 1-D tables along each axis, indexed by half-coordinate h in [0,2n_a] */
  double* pml_b[3];
  double* pml_c[3];
  int cpml_on;
  /* split-field layer (pml_kind DTM_PML_SPLIT): k1, k2 per half-coordinate,
   * and the part of each component along its second derivative axis */
  int split_on;
  double* sk1[3];
  double* sk2[3];
  real* spE[3];
  real* spH[3];
  uint8_t* mat[3];
  double* design_eps; /* per cell, 0 = not overridden */
  /* ADE state (per E component, per cell) */
  int any_disp, maxpoles;
  real* Em[3];
  real* Jp[3][DTM_MAX_POLES];
  real* Jm[3][DTM_MAX_POLES];
  /* gyrotropic plasma current, co-located with E */
  int any_gyro;
  real* Jg[3];
  /* coefficient tables per material */
  double inv_eps[DTM_MAX_MATS];
  int is_ade[DTM_MAX_MATS];
  double kE[DTM_MAX_MATS][3];
  double kEJ[DTM_MAX_MATS][DTM_MAX_POLES][2];
  double kJ[DTM_MAX_MATS][DTM_MAX_POLES][5];
  double gyA[DTM_MAX_MATS][9]; /* J^{n+1/2} = A J^{n-1/2} + dt wp^2 Minv E^n */
  double gyB[DTM_MAX_MATS][9];
  int is_gyro[DTM_MAX_MATS];
  /* probes and DFT */
  int nprobe;
  orc_probe probe[MAXPROBES];
  double* series; /* [cap][nprobe] */
  long long series_cap, series_len;
  int nfreq;
  double* omegas;
  double* dft; /* [nprobe][nfreq][2] */
  long long step;
} orc_sim;

static size_t IDX(const orc_sim* s, int i, int j, int k) {
  return ((size_t)i * s->ny + j) * s->nz + k;
}

/* Field value with the boundary rule (zero outside, or periodic wrap). */
static real getf(const orc_sim* s, const real* F, int i, int j, int k) {
  const int* bc = s->cfg.bc;
  if (i < 0 || i >= s->nx) {
    if (bc[0] != DTM_BC_PERIODIC) return 0;
    i = ((i % s->nx) + s->nx) % s->nx;
  }
  if (j < 0 || j >= s->ny) {
    if (bc[1] != DTM_BC_PERIODIC) return 0;
    j = ((j % s->ny) + s->ny) % s->ny;
  }
  if (k < 0 || k >= s->nz) {
    if (bc[2] != DTM_BC_PERIODIC) return 0;
    k = ((k % s->nz) + s->nz) % s->nz;
  }
  return F[IDX(s, i, j, k)];
}

/* ------------------------------------------------------ source functions */

/* Time function f(phase) and its first two phase-derivatives.
 * RAISED_COS: sin(p)*0.5*(1-cos(kEnv p)) on 0<=kEnv p<=2pi (signal.hpp:71-77).
 * GAUSS_SINE: exp(-((p/w-t0)/tw)^2) sin(p).
 * GAUSS: exp(-((p/w-t0)/tw)^2), the reference's EnvelopeTime pulse
 * (signal.hpp:55-57; derivatives per phase, EnvelopeDiv1/2 divided by w, w^2). */
static void wave3(int waveform, double p, double w, double kEnv, double t0, double tw,
                  double* f0, double* f1, double* f2) {
  if (waveform == DTM_WAVE_RAISED_COS) {
    double kp = p * kEnv;
    if (kp < 0 || kp > 2 * M_PI) {
      *f0 = *f1 = *f2 = 0;
      return;
    }
    double sp = sin(p), cp = cos(p), sk = sin(kp), ck = cos(kp);
    *f0 = sp * 0.5 * (1 - ck);
    *f1 = 0.5 * cp * (1 - ck) + 0.5 * kEnv * sp * sk;
    *f2 = -0.5 * sp * (1 - ck) + kEnv * cp * sk + 0.5 * kEnv * kEnv * sp * ck;
  } else {
    double a = 1.0 / (w * tw);
    double u = (p / w - t0) / tw;
    double g = exp(-u * u);
    double g1 = -2 * u * a * g;
    double g2 = (4 * u * u * a * a - 2 * a * a) * g;
    if (waveform == DTM_WAVE_GAUSS) { /* EnvelopeTime and its derivatives (signal.hpp:55-57), per phase */
      *f0 = g;
      *f1 = g1;
      *f2 = g2;
      return;
    }
    double sp = sin(p), cp = cos(p);
    *f0 = g * sp;
    *f1 = g1 * sp + g * cp;
    *f2 = g2 * sp + 2 * g1 * cp - g * sp;
  }
}

static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Incident field of component comp (DTM_EX..DTM_HZ) at physical (x,y,z), time t.
 * Dipole: signal.hpp:58-96.  Plane wave: signal.hpp:17-49 (no y-waist/z-cut). */
double orc_incident(const dtm_tfsf* src, int comp, double x, double y, double z, double t) {
  if (src->kind == DTM_TFSF_DIPOLE) {
    double xd = 0.5 * (src->box[0] + src->box[1]);
    double yd = 0.5 * (src->box[2] + src->box[3]);
    double zd = 0.5 * (src->box[4] + src->box[5]);
    double n[3] = {x - xd, y - yd, z - zd};
    double R = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double c = src->cSrc, w = src->w;
    double phase = w * t - w / c * R;
    double f0, f1, f2;
    wave3(src->waveform, phase, w, src->kEnv, src->t0, src->tw, &f0, &f1, &f2);
    /* dipole axis u = (sin th cos ph, sin th sin ph, cos th); th = 0 is the
     * reference's z dipole (signal.hpp:70-72) */
    double u[3] = {sin(src->theta) * cos(src->phi), sin(src->theta) * sin(src->phi), cos(src->theta)};
    double d[3], ds[3], dss[3];
    for (int q = 0; q < 3; q++) {
      d[q] = src->amp * f0 * u[q];
      ds[q] = src->amp * w * f1 * u[q];
      dss[q] = src->amp * w * w * f2 * u[q];
    }
    n[0] /= R; n[1] /= R; n[2] /= R;
    double t1[3], t2[3];
    if (comp < 3) {
      cross3(dss, n, t1);
      cross3(t1, n, t2);
      double dsn = ds[0] * n[0] + ds[1] * n[1] + ds[2] * n[2];
      double dn = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
      int a = comp;
      return c * (t2[a] / (c * c * R) + (3 * dsn * n[a] - ds[a]) / (c * R * R) +
                  (3 * dn * n[a] - d[a]) / (R * R * R));
    } else {
      int a = comp - 3;
      cross3(dss, n, t1);
      cross3(ds, n, t2);
      return t1[a] / (c * c * R) + t2[a] / (c * R * R);
    }
  } else if (src->kind == DTM_TFSF_PLANE) {
    double k = src->w / src->cSrc;
    double st = sin(src->theta), ct = cos(src->theta), sph = sin(src->phi), cph = cos(src->phi);
    double kx = k * st * cph, ky = k * st * sph, kz = k * ct;
    double phase = src->w * t - kx * (x - src->pos[0]) - ky * (y - src->pos[1]) - kz * (z - src->pos[2]);
    double f0, f1, f2;
    wave3(src->waveform, phase, src->w, src->kEnv, src->t0, src->tw, &f0, &f1, &f2);
    double s = src->amp * f0;
    double cps = cos(src->psi), sps = sin(src->psi);
    double Ephi = src->Imp * s * cps, Eth = src->Imp * s * sps;
    double Hth = -s * cps, Hphi = s * sps;
    switch (comp) {
      case DTM_EX: return -Ephi * sph + Eth * ct * cph;
      case DTM_EY: return Ephi * cph + Eth * ct * sph;
      case DTM_EZ: return -Eth * st;
      case DTM_HX: return -Hphi * sph + Hth * ct * cph;
      case DTM_HY: return Hphi * cph + Hth * ct * sph;
      case DTM_HZ: return -Hth * st;
    }
  }
  return 0;
}

/* Soft-source / waveform value at time t (phase = w t). */
double orc_wave(int waveform, double w, double kEnv, double t0, double tw, double t) {
  double f0, f1, f2;
  wave3(waveform, w * t, w, kEnv, t0, tw, &f0, &f1, &f2);
  return f0;
}

/* Half-coordinate of component comp in cell (i,j,k), see the lattice note in
 * include/dtmaxwell4.h (genAlg.py:24 typing). */
static void halfc(int comp, int i, int j, int k, int* h) {
  static const int off[6][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  h[0] = 2 * i + off[comp][0];
  h[1] = 2 * j + off[comp][1];
  h[2] = 2 * k + off[comp][2];
}

/* Component living at half-coordinates h (typing of genAlg.py:24). */
static int comp_at(const int* h) {
  int o0 = h[0] & 1, o1 = h[1] & 1, o2 = h[2] & 1;
  int no = o0 + o1 + o2;
  if (no == 2) return o0 == 0 ? DTM_EX : (o1 == 0 ? DTM_EY : DTM_EZ);
  if (no == 1) return o0 ? DTM_HX : (o1 ? DTM_HY : DTM_HZ);
  return -1;
}

/* inSF (signal.hpp:153-160): strict inside test on the physical position,
 * inverted when the scattered field lives outside the box. */
static int in_sf(const orc_sim* s, const int* h) {
  const dtm_tfsf* t = &s->cfg.tfsf;
  double x = h[0] * 0.5 * s->cfg.dx, y = h[1] * 0.5 * s->cfg.dy, z = h[2] * 0.5 * s->cfg.dz;
  int inside = x > t->box[0] && x < t->box[1] && y > t->box[2] && y < t->box[3] && z > t->box[4] &&
               z < t->box[5];
  return t->sf_inside ? inside : !inside;
}

static int tfsf_active(const orc_sim* s, long long n) {
  const dtm_config* c = &s->cfg;
  if (c->tfsf.kind == DTM_TFSF_NONE) return 0;
  int nt = c->ntime > 0 ? c->ntime : 1;
  long long istep = n / nt;
  return (double)((istep + 1) * nt) * c->dt < c->tfsf.tStop;
}

/* ------------------------------------------------------- CPML (A4) tables */

/* sigma u half-cells in from the outer face of a layer of p cells.
 * profile 0: sigma_max = V_max ln(1/R) (m+1)/(2 L) (DTmxw.cu:225), grading
 * sigma_max (depth/L)^m over the whole layer (DTmxw.cu:226-228).
 * profile 1: PMLgamma_func(i - 3, p' - 3) of DTmxw.cu:219-228 as
 * setPMLcoeffs calls it for x (DTmxw.cu:252), with i = u and p' = 2p half-
 * cells: i - 3 clamped at 0, zero from i - 3 >= N = p' - 6 on, sigma_max =
 * V ln(1/R) (m+1) / (2 N d/2) and sigma_max (1 - (i-3)/N)^m. */
static double orc_layer_sigma(int u, int p, double d, double m, double R, double vmax, int profile) {
  if (profile == DTM_PML_PROFILE_REFERENCE) {
    double N = 2.0 * p - 6;
    double a = u - 3 < 0 ? 0 : u - 3;
    if (a >= N) return 0;
    double smax = vmax * log(1.0 / R) * (m + 1) / (2 * (N * d * 0.5));
    double x = pow(smax, 1.0 / m) * (1 - a / N);
    return pow(x, m);
  }
  double L = p * d;
  double smax = vmax * log(1.0 / R) * (m + 1) / (2 * L);
  double depth = (double)(2 * p - u) / (2.0 * p);
  return smax * pow(depth, m);
}

/* sigma at half-coordinate h along an axis of n cells with thickness
 * plo/phi cells */
double orc_pml_sigma(int h, int n, int plo, int phi, double d, double m, double R, double vmax, int profile) {
  if (plo > 0 && h < 2 * plo) return orc_layer_sigma(h, plo, d, m, R, vmax, profile);
  if (phi > 0 && h > 2 * (n - phi)) return orc_layer_sigma(2 * n - h, phi, d, m, R, vmax, profile);
  return 0;
}

static void build_cpml(orc_sim* s) {
  const dtm_config* c = &s->cfg;
  int nn[3] = {s->nx, s->ny, s->nz};
  double dd[3] = {c->dx, c->dy, c->dz};
  const int split = c->pml_kind == DTM_PML_SPLIT;
  s->cpml_on = 0;
  s->split_on = 0;
  for (int a = 0; a < 3; a++) {
    s->pml_b[a] = (double*)calloc(2 * nn[a] + 1, sizeof(double));
    s->pml_c[a] = (double*)calloc(2 * nn[a] + 1, sizeof(double));
    s->sk1[a] = (double*)calloc(2 * nn[a] + 1, sizeof(double));
    s->sk2[a] = (double*)calloc(2 * nn[a] + 1, sizeof(double));
    for (int h = 0; h <= 2 * nn[a]; h++) {
      double sg = orc_pml_sigma(h, nn[a], c->npml[2 * a], c->npml[2 * a + 1], dd[a], c->pml_order,
                                c->pml_R, c->pml_vmax, c->pml_profile);
      /* split-field coefficients, setPMLcoeffs (DTmxw.cu:250-265) */
      s->sk2[a][h] = 1.0 / (1.0 + 0.5 * c->dt * sg);
      s->sk1[a][h] = 2.0 * s->sk2[a][h] - 1.0;
      if (split) {
        s->pml_b[a][h] = 1.0;
        s->pml_c[a][h] = 0.0;
        if (sg > 0) s->split_on = 1;
        continue;
      }
      double b = exp(-sg * c->dt);
      s->pml_b[a][h] = b;
      s->pml_c[a][h] = b - 1.0;
      if (sg > 0) s->cpml_on = 1;
    }
  }
}

/* -------------------------------------------------------- ADE (A7) tables */

/* DispStruct::set (dispersion.cpp:41-82): kE = {C1, C2, C3/dt},
 * kEJ = {C3(1+alpha)/2, C3 xi/2}, kJ = {alpha, xi, zeta+/dt, zeta-/dt, zeta/dt}. */
void orc_disp_coeffs(const dtm_material* m, double dt, double* kE, double* kEJ, double* kJ) {
  double zm = 0, zp = 0, z0 = 0;
  double al[DTM_MAX_POLES], xi[DTM_MAX_POLES], zetp[DTM_MAX_POLES], zetm[DTM_MAX_POLES],
      zet[DTM_MAX_POLES];
  for (int i = 0; i < m->npoles; i++) {
    const LorenzDisp* L = &m->pole[i];
    double den = L->b1 * dt + 2;
    al[i] = (4 - 2 * L->b0 * dt * dt) / den;
    xi[i] = (L->b1 * dt - 2) / den;
    zetp[i] = (L->a0 * dt * dt + 2 * L->a1 * dt) / den;
    zetm[i] = (-L->a0 * dt * dt + 2 * L->a1 * dt) / den;
    zet[i] = -(4 * L->a1 * dt) / den;
    zm += zetm[i];
    zp += zetp[i];
    z0 += zet[i];
  }
  double D = 2 * m->eps_inf + m->sigma * dt + zp;
  double C1 = -zm / D;
  double C2 = (2 * m->eps_inf - m->sigma * dt - z0) / D;
  double C3 = 2 * dt / D;
  kE[0] = C1;
  kE[1] = C2;
  kE[2] = C3 / dt;
  for (int i = 0; i < m->npoles; i++) {
    kEJ[2 * i + 0] = C3 * 0.5 * (1 + al[i]);
    kEJ[2 * i + 1] = C3 * 0.5 * xi[i];
    kJ[5 * i + 0] = al[i];
    kJ[5 * i + 1] = xi[i];
    kJ[5 * i + 2] = zetp[i] / dt;
    kJ[5 * i + 3] = zetm[i] / dt;
    kJ[5 * i + 4] = zet[i] / dt;
  }
}

/* Magnetised plasma current J (co-located with E, centred at half steps):
 * (1+nu dt/2) J^{n+1/2} - (dt/2) wc x J^{n+1/2}
 *   = (1-nu dt/2) J^{n-1/2} + (dt/2) wc x J^{n-1/2} + dt wp^2 E^n.
 * Returns A = M^-1 N and B = dt wp^2 M^-1 (row-major 3x3). */
void orc_gyro_coeffs(const dtm_material* m, double dt, double* A, double* B) {
  double w[3] = {m->gyro_wc[0], m->gyro_wc[1], m->gyro_wc[2]};
  /* W v = w x v */
  double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double M[9], N[9];
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) {
      double I = r == q ? 1.0 : 0.0;
      M[3 * r + q] = I * (1 + 0.5 * m->gyro_nu * dt) - 0.5 * dt * W[3 * r + q];
      N[3 * r + q] = I * (1 - 0.5 * m->gyro_nu * dt) + 0.5 * dt * W[3 * r + q];
    }
  double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
               M[2] * (M[3] * M[7] - M[4] * M[6]);
  double Mi[9];
  Mi[0] = (M[4] * M[8] - M[5] * M[7]) / det;
  Mi[1] = (M[2] * M[7] - M[1] * M[8]) / det;
  Mi[2] = (M[1] * M[5] - M[2] * M[4]) / det;
  Mi[3] = (M[5] * M[6] - M[3] * M[8]) / det;
  Mi[4] = (M[0] * M[8] - M[2] * M[6]) / det;
  Mi[5] = (M[2] * M[3] - M[0] * M[5]) / det;
  Mi[6] = (M[3] * M[7] - M[4] * M[6]) / det;
  Mi[7] = (M[1] * M[6] - M[0] * M[7]) / det;
  Mi[8] = (M[0] * M[4] - M[1] * M[3]) / det;
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) {
      double a = 0;
      for (int t = 0; t < 3; t++) a += Mi[3 * r + t] * N[3 * t + q];
      A[3 * r + q] = a;
      B[3 * r + q] = dt * m->gyro_wp * m->gyro_wp * Mi[3 * r + q];
    }
}

/* CoffStruct::set_eps restatement (defs.h:96-117), doubles out:
 * out[0..8] = deps matrix (XX,XY,XZ,YX,YY,YZ,ZX,ZY,ZZ with the reference's
 * off-diagonal naming coffAniso), out[9..17] = gyration matrix. */
void orc_set_eps(double eps, double gx, double gy, double gz, double w, double dt, double* out) {
  double ex = eps, ey = eps, ez = eps;
  gx *= w * dt;
  gy *= w * dt;
  gz *= w * dt;
  double D = 4 * ex * ey * ez + ex * gx * gx + ey * gy * gy + ez * gz * gz;
  out[0] = (gx * gx + 4 * ey * ez) / D;
  out[1] = (gx * gy + 2 * ez * gz) / D;
  out[2] = (gx * gy - 2 * ey * gy) / D;
  out[3] = (gx * gy - 2 * ez * gz) / D;
  out[4] = (gy * gy + 4 * ex * ez) / D;
  out[5] = (gy * gz + 2 * ex * gx) / D;
  out[6] = (gx * gz + 2 * ey * gy) / D;
  out[7] = (gy * gz - 2 * ex * gx) / D;
  out[8] = (gz * gz + 4 * ex * ey) / D;
  out[9] = (4 * ex * ey * ez + ex * gx * gx - ey * gy * gy - ez * gz * gz) / D;
  out[10] = (4 * ey * ez * gz + 2 * ey * gx * gy) / D;
  out[11] = (-4 * ey * ez * gy + 2 * ez * gx * gz) / D;
  out[12] = (-4 * ex * ez * gz + 2 * ex * gx * gy) / D;
  out[13] = (4 * ex * ey * ez + ey * gy * gy - ez * gz * gz - ex * gx * gx) / D;
  out[14] = (4 * ex * ez * gx + 2 * ez * gy * gz) / D;
  out[15] = (4 * ex * ey * gy + 2 * ex * gx * gz) / D;
  out[16] = (-4 * ex * ey * gx + 2 * ey * gy * gz) / D;
  out[17] = (4 * ex * ey * ez + ez * gz * gz - ex * gx * gx - ey * gy * gy) / D;
}

/* Purcell factor from E and dipole spectra (purcell.gp:2):
 * P = (A_E cos(phE) A_d sin(phd) - A_E sin(phE) A_d cos(phd)) w/2 / (A_d^2 w^4/3) + 1 */
double orc_purcell(double w, double AE, double phE, double Ad, double phd) {
  return (AE * cos(phE) * Ad * sin(phd) - AE * sin(phE) * Ad * cos(phd)) * w / 2 /
             (Ad * Ad * w * w * w * w / 3) +
         1;
}

/* --------------------------------------------------------------- lifecycle */

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n, sz);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz);
    abort();
  }
  return p;
}

orc_sim* orc_create(const dtm_config* cfg) {
  orc_sim* s = (orc_sim*)xcalloc(1, sizeof(orc_sim));
  s->cfg = *cfg;
  s->nx = cfg->nx;
  s->ny = cfg->ny;
  s->nz = cfg->nz;
  s->n = (size_t)s->nx * s->ny * s->nz;
  for (int c = 0; c < 3; c++) {
    s->E[c] = (real*)xcalloc(s->n, sizeof(real));
    s->H[c] = (real*)xcalloc(s->n, sizeof(real));
    s->mat[c] = (uint8_t*)xcalloc(s->n, 1);
  }
  build_cpml(s);
  if (s->split_on)
    for (int c = 0; c < 3; c++) {
      s->spE[c] = (real*)xcalloc(s->n, sizeof(real));
      s->spH[c] = (real*)xcalloc(s->n, sizeof(real));
    }
  if (s->cpml_on) {
    for (int c = 0; c < 3; c++)
      for (int a = 0; a < 3; a++)
        if (a != c) {
          s->psiE[c][a] = (real*)xcalloc(s->n, sizeof(real));
          s->psiH[c][a] = (real*)xcalloc(s->n, sizeof(real));
        }
  }
  s->any_disp = 0;
  s->any_gyro = 0;
  s->maxpoles = 0;
  for (int m = 0; m < cfg->nmat && m < DTM_MAX_MATS; m++) {
    const dtm_material* M = &cfg->mat[m];
    double eps = M->eps_inf > 0 ? M->eps_inf : 1.0;
    s->inv_eps[m] = 1.0 / eps;
    s->is_ade[m] = (M->npoles > 0 || M->sigma != 0) ? 1 : 0;
    s->is_gyro[m] = M->gyro ? 1 : 0;
    if (s->is_ade[m]) {
      s->any_disp = 1;
      if (M->npoles > s->maxpoles) s->maxpoles = M->npoles;
      orc_disp_coeffs(M, cfg->dt, s->kE[m], &s->kEJ[m][0][0], &s->kJ[m][0][0]);
    }
    if (s->is_gyro[m]) {
      s->any_gyro = 1;
      orc_gyro_coeffs(M, cfg->dt, s->gyA[m], s->gyB[m]);
    }
  }
  if (s->any_disp) {
    for (int c = 0; c < 3; c++) {
      s->Em[c] = (real*)xcalloc(s->n, sizeof(real));
      for (int p = 0; p < s->maxpoles; p++) {
        s->Jp[c][p] = (real*)xcalloc(s->n, sizeof(real));
        s->Jm[c][p] = (real*)xcalloc(s->n, sizeof(real));
      }
    }
  }
  if (s->any_gyro)
    for (int c = 0; c < 3; c++) s->Jg[c] = (real*)xcalloc(s->n, sizeof(real));
  return s;
}

void orc_destroy(orc_sim* s) {
  if (!s) return;
  for (int c = 0; c < 3; c++) {
    free(s->E[c]);
    free(s->H[c]);
    free(s->mat[c]);
    free(s->Em[c]);
    free(s->Jg[c]);
    for (int p = 0; p < DTM_MAX_POLES; p++) {
      free(s->Jp[c][p]);
      free(s->Jm[c][p]);
    }
    for (int a = 0; a < 3; a++) {
      free(s->psiE[c][a]);
      free(s->psiH[c][a]);
    }
    free(s->pml_b[c]);
    free(s->pml_c[c]);
    free(s->sk1[c]);
    free(s->sk2[c]);
    free(s->spE[c]);
    free(s->spH[c]);
  }
  free(s->design_eps);
  free(s->series);
  free(s->omegas);
  free(s->dft);
  free(s);
}

int orc_set_materials(orc_sim* s, const uint8_t* mx, const uint8_t* my, const uint8_t* mz) {
  const uint8_t* src[3] = {mx, my, mz};
  for (int c = 0; c < 3; c++) {
    if (src[c]) {
      for (size_t q = 0; q < s->n; q++)
        if (src[c][q] >= s->cfg.nmat) return DTM_ERR_ARG;
      memcpy(s->mat[c], src[c], s->n);
    } else {
      memset(s->mat[c], 0, s->n);
    }
  }
  return 0;
}

int orc_set_design_eps(orc_sim* s, int i0, int j0, int k0, int ni, int nj, int nk, const double* eps) {
  if (!s->design_eps) s->design_eps = (double*)xcalloc(s->n, sizeof(double));
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++)
        s->design_eps[IDX(s, i0 + i, j0 + j, k0 + k)] = eps[((size_t)i * nj + j) * nk + k];
  return 0;
}

int orc_add_probe(orc_sim* s, int kind, int comp, int i, int j, int k) {
  if (s->nprobe >= MAXPROBES || s->step > 0) return DTM_ERR_LIMIT;
  orc_probe* p = &s->probe[s->nprobe++];
  p->kind = kind;
  p->comp = comp;
  p->i = i;
  p->j = j;
  p->k = k;
  return s->nprobe - 1;
}

int orc_set_dft(orc_sim* s, int nfreq, const double* om) {
  free(s->omegas);
  free(s->dft);
  s->nfreq = nfreq;
  s->omegas = (double*)xcalloc(nfreq > 0 ? nfreq : 1, sizeof(double));
  memcpy(s->omegas, om, sizeof(double) * nfreq);
  s->dft = (double*)xcalloc((size_t)(s->nprobe > 0 ? s->nprobe : 1) * (nfreq > 0 ? nfreq : 1) * 2, sizeof(double));
  return 0;
}

/* --------------------------------------------------------------- stepping */

/* (2,4) difference of genAlg.py:196-198 with the sign folded in:
 * forward (E update, H samples at offsets -1..+2) and backward (H update). */
#define FWD(f_m1, f_0, f_p1, f_p2) (27.0 * ((f_p1) - (f_0)) - ((f_p2) - (f_m1)))
#define BWD(f_m2, f_m1, f_0, f_p1) (27.0 * ((f_0) - (f_m1)) - ((f_p1) - (f_m2)))

/* Value of F at a stencil neighbour (offset o along axis a) of cell (i,j,k). */
static real nb(const orc_sim* s, const real* F, int i, int j, int k, int a, int o) {
  if (a == 0) return getf(s, F, i + o, j, k);
  if (a == 1) return getf(s, F, i, j + o, k);
  return getf(s, F, i, j, k + o);
}

/* TF/SF correction for stencil neighbour q of point p (genAlg.py:219-221). */
static double tfsf_corr(const orc_sim* s, const int* hp, int psf, const int* hq, double t) {
  int qsf = in_sf(s, hq);
  if (qsf == psf) return 0;
  int qc = comp_at(hq);
  double v = orc_incident(&s->cfg.tfsf, qc, hq[0] * 0.5 * s->cfg.dx, hq[1] * 0.5 * s->cfg.dy,
                          hq[2] * 0.5 * s->cfg.dz, t);
  (void)hp;
  return (qsf && !psf) ? v : -v;
}

/* Derivative of field F along axis a for component `comp` at cell (i,j,k),
 * scaled by 1/(24 d_a), with TF/SF corrections when active. */
static double deriv(const orc_sim* s, const real* F, int comp, int i, int j, int k, int a,
                    int tfsf, double tinc) {
  const double dd[3] = {s->cfg.dx, s->cfg.dy, s->cfg.dz};
  int isE = comp < 3;
  /* offsets of the 4 samples, in the order (m2|m1, m1|0, 0|p1, p1|p2) */
  int o0 = isE ? -1 : -2, o1 = isE ? 0 : -1, o2 = isE ? 1 : 0, o3 = isE ? 2 : 1;
  double f[4] = {nb(s, F, i, j, k, a, o0), nb(s, F, i, j, k, a, o1), nb(s, F, i, j, k, a, o2),
                 nb(s, F, i, j, k, a, o3)};
  if (tfsf) {
    int hp[3];
    halfc(comp, i, j, k, hp);
    int psf = in_sf(s, hp);
    /* neighbours at half-offsets -3,-1,+1,+3 correspond to f[0..3] */
    static const int ho[4] = {-3, -1, 1, 3};
    for (int q = 0; q < 4; q++) {
      int hq[3] = {hp[0], hp[1], hp[2]};
      hq[a] += ho[q];
      f[q] += tfsf_corr(s, hp, psf, hq, tinc);
    }
  }
  double D = isE ? FWD(f[0], f[1], f[2], f[3]) : BWD(f[0], f[1], f[2], f[3]);
  return D / (24.0 * dd[a]);
}

/* The two derivative terms of curl for component c (0..2):
 * curl_c = D_{a1} F_{b1} - D_{a2} F_{b2}; (a1,b1,a2,b2) below. */
static const int CURL[3][4] = {
    {1, 2, 2, 1}, /* x: D_y Fz - D_z Fy */
    {2, 0, 0, 2}, /* y: D_z Fx - D_x Fz */
    {0, 1, 1, 0}, /* z: D_x Fy - D_y Fx */
};

static double cell_inv_eps(const orc_sim* s, int c, size_t q, int* mat_out) {
  int m = s->mat[c][q];
  *mat_out = m;
  if (s->design_eps && s->design_eps[q] > 0) return 1.0 / s->design_eps[q];
  return s->inv_eps[m];
}

/* Face bands of the TF/SF box: cells within 3 cells of a face plane along
 * each axis, and the box extent (+3 cells) along each axis. */
static int near_band(double pos, double lo, double hi, double d) {
  return fabs(pos - lo) <= 3 * d || fabs(pos - hi) <= 3 * d;
}
static int in_ext(double pos, double lo, double hi, double d) { return pos >= lo - 3 * d && pos <= hi + 3 * d; }

/* Rows (i,j) whose whole k-range can use the unchecked stencil: all four
 * samples along x and y exist, and no x/y TF/SF face is near.  A row that
 * crosses the box's z faces is still fast; the cells near those faces take
 * the general path (tfsf_zcell). */
static int row_fast(const orc_sim* s, int i, int j, int tf) {
  if (s->split_on) return 0; /* the split-field layer is on the general path only */
  if (s->cfg.bc[0] == DTM_BC_PERIODIC || s->cfg.bc[1] == DTM_BC_PERIODIC || s->cfg.bc[2] == DTM_BC_PERIODIC)
    return 0;
  if (i < 2 || i > s->nx - 3 || j < 2 || j > s->ny - 3 || s->nz < 5) return 0;
  if (tf) {
    const dtm_tfsf* t = &s->cfg.tfsf;
    double x = (i + 0.5) * s->cfg.dx, y = (j + 0.5) * s->cfg.dy;
    int inx = in_ext(x, t->box[0], t->box[1], s->cfg.dx), iny = in_ext(y, t->box[2], t->box[3], s->cfg.dy);
    if (near_band(x, t->box[0], t->box[1], s->cfg.dx) && iny) return 0;
    if (near_band(y, t->box[2], t->box[3], s->cfg.dy) && inx) return 0;
  }
  return 1;
}

/* Cell k of a fast row that must take the general path for TF/SF. */
static int tfsf_zcell(const orc_sim* s, int i, int j, int k) {
  const dtm_tfsf* t = &s->cfg.tfsf;
  double x = (i + 0.5) * s->cfg.dx, y = (j + 0.5) * s->cfg.dy, z = (k + 0.5) * s->cfg.dz;
  return in_ext(x, t->box[0], t->box[1], s->cfg.dx) && in_ext(y, t->box[2], t->box[3], s->cfg.dy) &&
         near_band(z, t->box[4], t->box[5], s->cfg.dz);
}

/* One fast H row: cells k in [2, nz-2) of row (i,j). */
static void h_cell_general(orc_sim* s, int i, int j, int k, int tf, double tinc);
static void h_row_fast(orc_sim* s, int i, int j, int tf, double tinc) {
  const size_t sx = (size_t)s->ny * s->nz, sy = s->nz;
  const real dt = (real)s->cfg.dt;
  const real rx = (real)(1.0 / (24.0 * s->cfg.dx)), ry = (real)(1.0 / (24.0 * s->cfg.dy)),
             rz = (real)(1.0 / (24.0 * s->cfg.dz));
  const real* Ex = s->E[0];
  const real* Ey = s->E[1];
  const real* Ez = s->E[2];
  real* Hx = s->H[0];
  real* Hy = s->H[1];
  real* Hz = s->H[2];
  const real c27 = (real)27;
  const size_t base = IDX(s, i, j, 0);
  const int hx_e = 2 * i, hy_e = 2 * j; /* even half-coords of H along x, y */
  const int on = s->cpml_on;
  const double bxv = on ? s->pml_c[0][hx_e] : 0, byv = on ? s->pml_c[1][hy_e] : 0;
  for (int k = 2; k < s->nz - 1; k++) {
    size_t q = base + k;
    if (tf && tfsf_zcell(s, i, j, k)) {
      h_cell_general(s, i, j, k, tf, tinc);
      continue;
    }
    real dyEz = ry * (c27 * (Ez[q] - Ez[q - sy]) - (Ez[q + sy] - Ez[q - 2 * sy]));
    real dzEy = rz * (c27 * (Ey[q] - Ey[q - 1]) - (Ey[q + 1] - Ey[q - 2]));
    real dzEx = rz * (c27 * (Ex[q] - Ex[q - 1]) - (Ex[q + 1] - Ex[q - 2]));
    real dxEz = rx * (c27 * (Ez[q] - Ez[q - sx]) - (Ez[q + sx] - Ez[q - 2 * sx]));
    real dxEy = rx * (c27 * (Ey[q] - Ey[q - sx]) - (Ey[q + sx] - Ey[q - 2 * sx]));
    real dyEx = ry * (c27 * (Ex[q] - Ex[q - sy]) - (Ex[q + sy] - Ex[q - 2 * sy]));
    if (on) {
      const int hz = 2 * k;
      double cz = s->pml_c[2][hz];
      if (byv != 0) {
        real* p = s->psiH[0][1]; p[q] = (real)(s->pml_b[1][hy_e] * p[q] + byv * dyEz); dyEz += p[q];
        p = s->psiH[2][1]; p[q] = (real)(s->pml_b[1][hy_e] * p[q] + byv * dyEx); dyEx += p[q];
      }
      if (cz != 0) {
        real* p = s->psiH[0][2]; p[q] = (real)(s->pml_b[2][hz] * p[q] + cz * dzEy); dzEy += p[q];
        p = s->psiH[1][2]; p[q] = (real)(s->pml_b[2][hz] * p[q] + cz * dzEx); dzEx += p[q];
      }
      if (bxv != 0) {
        real* p = s->psiH[1][0]; p[q] = (real)(s->pml_b[0][hx_e] * p[q] + bxv * dxEz); dxEz += p[q];
        p = s->psiH[2][0]; p[q] = (real)(s->pml_b[0][hx_e] * p[q] + bxv * dxEy); dxEy += p[q];
      }
    }
    Hx[q] -= dt * (dyEz - dzEy);
    Hy[q] -= dt * (dzEx - dxEz);
    Hz[q] -= dt * (dxEy - dyEx);
  }
}

static void h_cell_general(orc_sim* s, int i, int j, int k, int tf, double tinc) {
  size_t q = IDX(s, i, j, k);
  for (int c = 0; c < 3; c++) {
    int comp = 3 + c;
    int a1 = CURL[c][0], b1 = CURL[c][1], a2 = CURL[c][2], b2 = CURL[c][3];
    double d1 = deriv(s, s->E[b1], comp, i, j, k, a1, tf, tinc);
    double d2 = deriv(s, s->E[b2], comp, i, j, k, a2, tf, tinc);
    if (s->cpml_on) {
      int h[3];
      halfc(comp, i, j, k, h);
      double cc1 = s->pml_c[a1][h[a1]], cc2 = s->pml_c[a2][h[a2]];
      if (cc1 != 0) {
        real* ps = s->psiH[c][a1];
        ps[q] = (real)(s->pml_b[a1][h[a1]] * ps[q] + cc1 * d1);
        d1 += ps[q];
      }
      if (cc2 != 0) {
        real* ps = s->psiH[c][a2];
        ps[q] = (real)(s->pml_b[a2][h[a2]] * ps[q] + cc2 * d2);
        d2 += ps[q];
      }
    }
    if (s->split_on) {
      /* genAlg.py:117-148: H = P1 + P2, P_a = k1_a P_a -+ k2_a dt dif_a */
      int h[3];
      halfc(comp, i, j, k, h);
      double k11 = s->sk1[a1][h[a1]], k21 = s->sk2[a1][h[a1]];
      double k12 = s->sk1[a2][h[a2]], k22 = s->sk2[a2][h[a2]];
      if (k21 != 1.0 || k22 != 1.0) {
        double p2 = s->spH[c][q], p1 = s->H[c][q] - p2;
        p1 = k11 * p1 - k21 * s->cfg.dt * d1;
        p2 = k12 * p2 + k22 * s->cfg.dt * d2;
        s->spH[c][q] = (real)p2;
        s->H[c][q] = (real)(p1 + p2);
        continue;
      }
    }
    s->H[c][q] = (real)(s->H[c][q] - s->cfg.dt * (d1 - d2));
  }
}

static void h_update(orc_sim* s) {
  long long n = s->step;
  int tf = tfsf_active(s, n);
  double tinc = (double)n * s->cfg.dt; /* E incident at t = n dt (genAlg.py:220, time=0) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->nx; i++)
    for (int j = 0; j < s->ny; j++) {
      if (row_fast(s, i, j, tf)) {
        h_cell_general(s, i, j, 0, tf, tinc);
        h_cell_general(s, i, j, 1, tf, tinc);
        h_row_fast(s, i, j, tf, tinc);
        h_cell_general(s, i, j, s->nz - 1, tf, tinc);
      } else {
        for (int k = 0; k < s->nz; k++) h_cell_general(s, i, j, k, tf, tinc);
      }
    }
}

static void e_cell_general(orc_sim* s, int i, int j, int k, int tf, double tinc, const double* Jsrc) {
  const dtm_config* cfg = &s->cfg;
  double dt = cfg->dt;
  size_t q = IDX(s, i, j, k);
  double curl[3], dd1[3], dd2[3];
  for (int c = 0; c < 3; c++) {
    int a1 = CURL[c][0], b1 = CURL[c][1], a2 = CURL[c][2], b2 = CURL[c][3];
    double d1 = deriv(s, s->H[b1], c, i, j, k, a1, tf, tinc);
    double d2 = deriv(s, s->H[b2], c, i, j, k, a2, tf, tinc);
    if (s->cpml_on) {
      int h[3];
      halfc(c, i, j, k, h);
      double cc1 = s->pml_c[a1][h[a1]], cc2 = s->pml_c[a2][h[a2]];
      if (cc1 != 0) {
        real* ps = s->psiE[c][a1];
        ps[q] = (real)(s->pml_b[a1][h[a1]] * ps[q] + cc1 * d1);
        d1 += ps[q];
      }
      if (cc2 != 0) {
        real* ps = s->psiE[c][a2];
        ps[q] = (real)(s->pml_b[a2][h[a2]] * ps[q] + cc2 * d2);
        d2 += ps[q];
      }
    }
    curl[c] = d1 - d2;
    dd1[c] = d1;
    dd2[c] = d2;
    for (int so = 0; so < cfg->nsoft; so++) {
      const dtm_soft_source* ss = &cfg->soft[so];
      if (ss->comp == c && ss->i == i && ss->j == j && ss->k == k) {
        curl[c] -= Jsrc[so];
        dd1[c] -= Jsrc[so];
      }
    }
  }
  /* gyrotropic current: J^{n+1/2} from J^{n-1/2} and E^n (co-located) */
  int mg = s->mat[0][q];
  if (s->any_gyro && s->is_gyro[mg]) {
    double J0[3] = {s->Jg[0][q], s->Jg[1][q], s->Jg[2][q]};
    double E0[3] = {s->E[0][q], s->E[1][q], s->E[2][q]};
    for (int r = 0; r < 3; r++) {
      double v = 0;
      for (int t = 0; t < 3; t++) v += s->gyA[mg][3 * r + t] * J0[t] + s->gyB[mg][3 * r + t] * E0[t];
      s->Jg[r][q] = (real)v;
    }
    for (int r = 0; r < 3; r++) {
      curl[r] -= s->Jg[r][q];
      dd1[r] -= s->Jg[r][q];
    }
  }
  for (int c = 0; c < 3; c++) {
    int m;
    double ie = cell_inv_eps(s, c, q, &m);
    if (s->any_disp && s->is_ade[m] && !(s->design_eps && s->design_eps[q] > 0)) {
      /* countDisp (dispersion.cpp:85-99), rotdt = dt*curl */
      double Ep = s->E[c][q], Em = s->Em[c][q];
      double rotdt = dt * curl[c];
      double sum = 0;
      int np = cfg->mat[m].npoles;
      for (int p = 0; p < np; p++)
        sum += s->kEJ[m][p][0] * s->Jp[c][p][q] + s->kEJ[m][p][1] * s->Jm[c][p][q];
      double En = s->kE[m][0] * Em + s->kE[m][1] * Ep + s->kE[m][2] * rotdt - sum;
      for (int p = 0; p < np; p++) {
        double jp = s->Jp[c][p][q], jm = s->Jm[c][p][q];
        double jn = s->kJ[m][p][0] * jp + s->kJ[m][p][1] * jm + s->kJ[m][p][2] * En +
                    s->kJ[m][p][3] * Em + s->kJ[m][p][4] * Ep;
        s->Jm[c][p][q] = (real)jp;
        s->Jp[c][p][q] = (real)jn;
      }
      s->Em[c][q] = (real)Ep;
      s->E[c][q] = (real)En;
    } else {
      if (s->split_on) {
        /* genAlg.py:117-148 on E: E = P1 + P2, P_a = k1_a P_a +- k2_a dt/eps dif_a */
        int h[3];
        int a1 = CURL[c][0], a2 = CURL[c][2];
        halfc(c, i, j, k, h);
        double k11 = s->sk1[a1][h[a1]], k21 = s->sk2[a1][h[a1]];
        double k12 = s->sk1[a2][h[a2]], k22 = s->sk2[a2][h[a2]];
        if (k21 != 1.0 || k22 != 1.0) {
          double p2 = s->spE[c][q], p1 = s->E[c][q] - p2;
          p1 = k11 * p1 + k21 * dt * ie * dd1[c];
          p2 = k12 * p2 - k22 * dt * ie * dd2[c];
          s->spE[c][q] = (real)p2;
          s->E[c][q] = (real)(p1 + p2);
          continue;
        }
      }
      s->E[c][q] = (real)(s->E[c][q] + dt * ie * curl[c]);
    }
  }
}

static int cell_special(const orc_sim* s, size_t q) {
  if (s->design_eps && s->design_eps[q] > 0) return 1;
  for (int c = 0; c < 3; c++) {
    int m = s->mat[c][q];
    if (s->is_ade[m] || s->is_gyro[m]) return 1;
  }
  return 0;
}

/* One fast E row: cells k in [1, nz-2) of row (i,j). */
static void e_row_fast(orc_sim* s, int i, int j, int tf, double tinc, const double* Jsrc) {
  const size_t sx = (size_t)s->ny * s->nz, sy = s->nz;
  const real dt = (real)s->cfg.dt;
  const real rx = (real)(1.0 / (24.0 * s->cfg.dx)), ry = (real)(1.0 / (24.0 * s->cfg.dy)),
             rz = (real)(1.0 / (24.0 * s->cfg.dz));
  const real* Hx = s->H[0];
  const real* Hy = s->H[1];
  const real* Hz = s->H[2];
  real* Ex = s->E[0];
  real* Ey = s->E[1];
  real* Ez = s->E[2];
  const real c27 = (real)27;
  const size_t base = IDX(s, i, j, 0);
  const int hx_o = 2 * i + 1, hy_o = 2 * j + 1; /* odd half-coords of E along x, y */
  const int on = s->cpml_on;
  const double cxv = on ? s->pml_c[0][hx_o] : 0, cyv = on ? s->pml_c[1][hy_o] : 0;
  int row_src = 0;
  for (int so = 0; so < s->cfg.nsoft; so++)
    if (s->cfg.soft[so].i == i && s->cfg.soft[so].j == j) row_src = 1;
  for (int k = 1; k < s->nz - 2; k++) {
    size_t q = base + k;
    if (row_src || cell_special(s, q) || (tf && tfsf_zcell(s, i, j, k))) {
      e_cell_general(s, i, j, k, tf, tinc, Jsrc);
      continue;
    }
    real dyHz = ry * (c27 * (Hz[q + sy] - Hz[q]) - (Hz[q + 2 * sy] - Hz[q - sy]));
    real dzHy = rz * (c27 * (Hy[q + 1] - Hy[q]) - (Hy[q + 2] - Hy[q - 1]));
    real dzHx = rz * (c27 * (Hx[q + 1] - Hx[q]) - (Hx[q + 2] - Hx[q - 1]));
    real dxHz = rx * (c27 * (Hz[q + sx] - Hz[q]) - (Hz[q + 2 * sx] - Hz[q - sx]));
    real dxHy = rx * (c27 * (Hy[q + sx] - Hy[q]) - (Hy[q + 2 * sx] - Hy[q - sx]));
    real dyHx = ry * (c27 * (Hx[q + sy] - Hx[q]) - (Hx[q + 2 * sy] - Hx[q - sy]));
    if (on) {
      const int hz = 2 * k + 1;
      double cz = s->pml_c[2][hz];
      if (cyv != 0) {
        real* p = s->psiE[0][1]; p[q] = (real)(s->pml_b[1][hy_o] * p[q] + cyv * dyHz); dyHz += p[q];
        p = s->psiE[2][1]; p[q] = (real)(s->pml_b[1][hy_o] * p[q] + cyv * dyHx); dyHx += p[q];
      }
      if (cz != 0) {
        real* p = s->psiE[0][2]; p[q] = (real)(s->pml_b[2][hz] * p[q] + cz * dzHy); dzHy += p[q];
        p = s->psiE[1][2]; p[q] = (real)(s->pml_b[2][hz] * p[q] + cz * dzHx); dzHx += p[q];
      }
      if (cxv != 0) {
        real* p = s->psiE[1][0]; p[q] = (real)(s->pml_b[0][hx_o] * p[q] + cxv * dxHz); dxHz += p[q];
        p = s->psiE[2][0]; p[q] = (real)(s->pml_b[0][hx_o] * p[q] + cxv * dxHy); dxHy += p[q];
      }
    }
    real ie0, ie1, ie2;
    if (s->design_eps == NULL) {
      ie0 = (real)s->inv_eps[s->mat[0][q]];
      ie1 = (real)s->inv_eps[s->mat[1][q]];
      ie2 = (real)s->inv_eps[s->mat[2][q]];
    } else {
      int m;
      ie0 = (real)cell_inv_eps(s, 0, q, &m);
      ie1 = (real)cell_inv_eps(s, 1, q, &m);
      ie2 = (real)cell_inv_eps(s, 2, q, &m);
    }
    Ex[q] += dt * ie0 * (dyHz - dzHy);
    Ey[q] += dt * ie1 * (dzHx - dxHz);
    Ez[q] += dt * ie2 * (dxHy - dyHx);
  }
}

static void e_update(orc_sim* s) {
  long long n = s->step;
  int tf = tfsf_active(s, n);
  double tinc = ((double)n + 0.5) * s->cfg.dt; /* H incident at t=(n+1/2)dt (genAlg.py:499) */
  const dtm_config* cfg = &s->cfg;
  /* soft current sources J(t_{n+1/2}) */
  double Jsrc[DTM_MAX_SOFT];
  for (int q = 0; q < cfg->nsoft; q++) {
    const dtm_soft_source* so = &cfg->soft[q];
    Jsrc[q] = so->amp * orc_wave(so->waveform, so->w, so->kEnv, so->t0, so->tw, tinc);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->nx; i++)
    for (int j = 0; j < s->ny; j++) {
      if (row_fast(s, i, j, tf)) {
        e_cell_general(s, i, j, 0, tf, tinc, Jsrc);
        e_row_fast(s, i, j, tf, tinc, Jsrc);
        e_cell_general(s, i, j, s->nz - 2, tf, tinc, Jsrc);
        e_cell_general(s, i, j, s->nz - 1, tf, tinc, Jsrc);
      } else {
        for (int k = 0; k < s->nz; k++) e_cell_general(s, i, j, k, tf, tinc, Jsrc);
      }
    }
}

/* Reference sensor value (sensor.h:64-71), Yee-node coordinates. */
static double sensor_value(const orc_sim* s, int comp, int x, int y, int z) {
  int zM = (z == 0) ? z + 1 : z - 1; /* sensor.h:64 */
  switch (comp) {
    case DTM_EX:
      return 0.25 * (getf(s, s->E[0], x, y - 1, zM) + getf(s, s->E[0], x, y - 1, z) +
                     getf(s, s->E[0], x, y, zM) + getf(s, s->E[0], x, y, z));
    case DTM_EY:
      return 0.25 * (getf(s, s->E[1], x - 1, y, zM) + getf(s, s->E[1], x - 1, y, z) +
                     getf(s, s->E[1], x, y, zM) + getf(s, s->E[1], x, y, z));
    case DTM_EZ:
      return 0.25 * (getf(s, s->E[2], x - 1, y - 1, z) + getf(s, s->E[2], x - 1, y, z) +
                     getf(s, s->E[2], x, y - 1, z) + getf(s, s->E[2], x, y, z));
    case DTM_HX:
      return 0.5 * (getf(s, s->H[0], x - 1, y, z) + getf(s, s->H[0], x, y, z));
    case DTM_HY:
      return 0.5 * (getf(s, s->H[1], x, y - 1, z) + getf(s, s->H[1], x, y, z));
    case DTM_HZ:
      return 0.5 * (getf(s, s->H[2], x, y, zM) + getf(s, s->H[2], x, y, z));
  }
  return 0;
}

static void record(orc_sim* s) {
  if (s->nprobe == 0) return;
  if (s->series_len >= s->series_cap) {
    long long cap = s->series_cap ? 2 * s->series_cap : 1024;
    s->series = (double*)realloc(s->series, sizeof(double) * cap * s->nprobe);
    s->series_cap = cap;
  }
  double* row = s->series + s->series_len * s->nprobe;
  long long n = s->step;
  for (int p = 0; p < s->nprobe; p++) {
    const orc_probe* pr = &s->probe[p];
    double v;
    if (pr->kind == 0)
      v = pr->comp < 3 ? s->E[pr->comp][IDX(s, pr->i, pr->j, pr->k)]
                       : s->H[pr->comp - 3][IDX(s, pr->i, pr->j, pr->k)];
    else
      v = sensor_value(s, pr->comp, pr->i, pr->j, pr->k);
    row[p] = v;
    if (s->nfreq > 0 && s->dft) {
      double t = pr->comp < 3 ? (n + 1.0) * s->cfg.dt : (n + 0.5) * s->cfg.dt;
      for (int f = 0; f < s->nfreq; f++) {
        double ph = s->omegas[f] * t;
        s->dft[((size_t)p * s->nfreq + f) * 2 + 0] += v * cos(ph) * s->cfg.dt;
        s->dft[((size_t)p * s->nfreq + f) * 2 + 1] += v * sin(ph) * s->cfg.dt;
      }
    }
  }
  s->series_len++;
}

int orc_step(orc_sim* s, int nsteps) {
  for (int it = 0; it < nsteps; it++) {
    h_update(s);
    e_update(s);
    record(s);
    s->step++;
  }
  return 0;
}

long long orc_time_step(const orc_sim* s) { return s->step; }

/* Current value of a reference sensor (kind 1) or raw probe (kind 0). */
double orc_probe_now(const orc_sim* s, int kind, int comp, int i, int j, int k) {
  if (kind == 0) return comp < 3 ? s->E[comp][IDX(s, i, j, k)] : s->H[comp - 3][IDX(s, i, j, k)];
  return sensor_value(s, comp, i, j, k);
}

int orc_field_read(const orc_sim* s, int comp, double* out) {
  const real* F = comp < 3 ? s->E[comp] : s->H[comp - 3];
  for (size_t q = 0; q < s->n; q++) out[q] = (double)F[q];
  return 0;
}

int orc_field_write(orc_sim* s, int comp, const double* in) {
  real* F = comp < 3 ? s->E[comp] : s->H[comp - 3];
  for (size_t q = 0; q < s->n; q++) F[q] = (real)in[q];
  return 0;
}

int orc_probe_read(const orc_sim* s, double* out, long long rows) {
  long long r = rows < s->series_len ? rows : s->series_len;
  memcpy(out, s->series, sizeof(double) * r * s->nprobe);
  return (int)r;
}

int orc_dft_read(const orc_sim* s, double* out) {
  if (!s->dft) return 0;
  memcpy(out, s->dft, sizeof(double) * s->nprobe * s->nfreq * 2);
  return 0;
}

int orc_nthreads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_real_size(void) { return (int)sizeof(real); }
