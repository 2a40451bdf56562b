// dtm_solver.cpp — host side of the C ABI (include/dtmaxwell4.h).
//
// One dtm_solver owns the device state of one grid on the HIP device that was
// current at dtm_create, and one HIP stream (default priority).  All kernel
// launches are asynchronous on that stream; reads synchronise it.
//
// Reference counterparts: GeoParamsHost::set (DTmxw.cu:278-487) for the
// checks and allocations, init_index / setPMLcoeffs (DTmxw.cu:250-265,
// 502-517) for the absorbing-layer tables, init_material (DTmxw.cu:604-642)
// and DispStruct::set (dispersion.cpp:41-82) for the coefficient tables,
// TFSFsrc (signal.hpp:99-160) for the source, Sensor (sensor.h) for probes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dtmaxwell4.h"
#include "dtm_internal.h"

using namespace dtm;

static thread_local std::string g_err;
static constexpr int kTableSteps = 32;  // plane-wave waveform tables per batch

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// DTM_ERR_DEVICE with the kernel and reason when a launch since the last
// call failed or was refused (see note_launch).
static int launch_result() {
  char m[256];
  if (take_launch_error(m, sizeof(m))) return fail(DTM_ERR_DEVICE, m);
  return 0;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return fail(DTM_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

struct dtm_solver {
  dtm_config cfg;
  int dev = 0;
  hipStream_t stream = nullptr;
  SideStream side{};  // the FP32 zone's column runs on side.st (set with the zone)
  Grid g{};
  Scales sc{};
  int store = 0;
  void* E[3] = {};   // current fields (E^n, H^{n-1/2} before a step)
  void* H[3] = {};
  void* E2[3] = {};  // second buffers of the fused (torre) step
  void* H2[3] = {};
  Fields<0> f32{}, f32n{};
  Fields<1> f16{}, f16n{};
  Cpml pm{};         // current psi
  Cpml pmn{};        // second psi buffers (torre step)
  bool torre = false; // fused single-pass step; false: two sweeps per step
  bool no_fold = false; // tabulated plane-wave TF/SF through k_tfsf instead of the sweeps
  float* fold_base = nullptr; // per-plane factors of a folded plane wave, kTableSteps x 2 half-steps
  // device copies of the fused step's pointer table, one per buffer parity
  // (the old / new buffers swap every step), and their host images
  TorreCold* cold_dev[2] = {};
  TorreCold cold_host[2];
  bool cold_ok[2] = {false, false};
  int cold_last = 1;
  std::vector<void*> allocs;  // everything to free
  std::vector<std::pair<void*, size_t>> state;  // field / psi buffers (zeroed by a reset)
  bool adjoint = false;  // stepping adjoint fields: no sources, no recording
  bool replay = false;   // recomputing a forward segment of the adjoint: no recording
  int adj_every = 0;     // adjoint: 0 = keep the design-box history, K > 0 = checkpoint every K steps
  Region whole{};      // the grid: one launch of the variant with every boundary rule
  // materials
  std::vector<uint8_t> cls_host;
  std::vector<ClassEntry> classes;
  std::vector<std::array<uint8_t, 3>> class_mats;
  uint8_t* cls_dev = nullptr;
  ClassEntry* table_dev = nullptr;
  Materials mt{};
  Ade ade{};
  AdeMat* ademats_dev = nullptr;
  GyroMat* gyromats_dev = nullptr;
  float* design_dev = nullptr;
  std::vector<float> design_host;
  // TF/SF
  Tfsf tf{};
  TfsfBands bands{};
  double* tab_base = nullptr;  // waveform tables of kTableSteps steps x 2 halves
  long long tab_n0 = -1;       // first step held in tab_base
  // probes
  std::vector<ProbeDesc> probes;
  ProbeDesc* probes_dev = nullptr;
  float* series_dev = nullptr;
  int series_cap = 0;
  int series_row = 0;
  std::vector<float> series_host;
  int nfreq = 0;
  double* omega_dev = nullptr;
  double* dft_dev = nullptr;
  double* energy_dev = nullptr;
  double* energy_part = nullptr;  // box energy: one FP64 partial per x plane
  int energy_part_cap = 0;
  // Poynting-flux boxes (series rows share the probe series' row counter)
  std::vector<FluxFace> flux_faces;
  int flux_snap_n = 0;
  FluxSet fs{};
  std::vector<double> flux_host;
  float* tmp_dev = nullptr;  // conversion buffer (FP16 field access)
  long long step = 0;
  bool started = false;
  // per-group kernel timing
  bool timing = false;
  std::vector<std::array<hipEvent_t, 2>> ev_pool;
  std::vector<int> ev_group;
  size_t ev_used = 0;
  double grp_ms[5] = {};
  long long grp_cnt[5] = {};
};

// floats of one half-step's fold factor table: header + nx planes, 8 each
static long long fold_floats(const dtm_solver* s) { return ((long long)s->cfg.nx + 1) * kFoldStride; }

template <typename T>
static int dalloc(dtm_solver* s, T** p, size_t bytes) {
  void* q = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(DTM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemsetAsync(q, 0, bytes, s->stream);
  if (e != hipSuccess) return fail(DTM_ERR_DEVICE, hipGetErrorString(e));
  s->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return 0;
}

// dalloc for buffers that hold time-stepped state.
template <typename T>
static int sdalloc(dtm_solver* s, T** p, size_t bytes) {
  if (int rc = dalloc(s, p, bytes)) return rc;
  s->state.emplace_back(static_cast<void*>(*p), bytes == 0 ? 16 : bytes);
  return 0;
}

static int zero_state(dtm_solver* s) {
  for (auto& b : s->state) HIPCHK(hipMemsetAsync(b.first, 0, b.second, s->stream));
  return 0;
}

extern "C" {

const char* dtm_last_error(void) { return g_err.c_str(); }
const char* dtm_version(void) { return "dtmaxwell4-amd 0.1 (gfx950)"; }

int dtm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void LorenzDisp_set(LorenzDisp* p, double Deps, double wp, double gp, double gp_) {
  p->a0 = Deps * wp * wp;
  p->a1 = Deps * gp_;
  p->b0 = wp * wp;
  p->b1 = 2 * gp;
  p->b2 = 1;
}
void LorenzDisp_setDrude(LorenzDisp* p, double Deps, double wp, double gp, double gp_) {
  p->a0 = Deps * wp * wp;
  p->a1 = Deps * gp_;
  p->b0 = 0.0;
  p->b1 = 2 * gp;
  p->b2 = 1;
}

// CoffStruct::set_eps (defs.h:96-117): Cayley pair of the static gyration.
void CoffStruct_set_eps(CoffStruct* c, double eps, int disp, double gx, double gy, double gz, double w, double dt) {
  c->deps = (float)(1. / eps);
  c->isDisp = disp;
  const double ex = eps, ey = eps, ez = eps;
  gx *= w * dt;
  gy *= w * dt;
  gz *= w * dt;
  const double D = 4 * ex * ey * ez + ex * gx * gx + ey * gy * gy + ez * gz * gz;
  c->depsXX = (float)((gx * gx + 4 * ey * ez) / D);
  c->coffAnisoXY = (float)((gx * gy + 2 * ez * gz) / D);
  c->coffAnisoXZ = (float)((gx * gy - 2 * ey * gy) / D);
  c->coffAnisoYX = (float)((gx * gy - 2 * ez * gz) / D);
  c->depsYY = (float)((gy * gy + 4 * ex * ez) / D);
  c->coffAnisoYZ = (float)((gy * gz + 2 * ex * gx) / D);
  c->coffAnisoZX = (float)((gx * gz + 2 * ey * gy) / D);
  c->coffAnisoZY = (float)((gy * gz - 2 * ex * gx) / D);
  c->depsZZ = (float)((gz * gz + 4 * ex * ey) / D);
  c->gyrationXX = (float)((4 * ex * ey * ez + ex * gx * gx - ey * gy * gy - ez * gz * gz) / D);
  c->gyrationXY = (float)((4 * ey * ez * gz + 2 * ey * gx * gy) / D);
  c->gyrationXZ = (float)((-4 * ey * ez * gy + 2 * ez * gx * gz) / D);
  c->gyrationYX = (float)((-4 * ex * ez * gz + 2 * ex * gx * gy) / D);
  c->gyrationYY = (float)((4 * ex * ey * ez + ey * gy * gy - ez * gz * gz - ex * gx * gx) / D);
  c->gyrationYZ = (float)((4 * ex * ez * gx + 2 * ez * gy * gz) / D);
  c->gyrationZX = (float)((4 * ex * ey * gy + 2 * ex * gx * gz) / D);
  c->gyrationZY = (float)((-4 * ex * ey * gx + 2 * ey * gy * gz) / D);
  c->gyrationZZ = (float)((4 * ex * ey * ez + ez * gz * gz - ex * gx * gx - ey * gy * gy) / D);
}

void dtm_config_defaults(dtm_config* c, int nx, int ny, int nz) {
  memset(c, 0, sizeof(*c));
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->dx = c->dy = c->dz = 0.005;  // py_consts.h:71
  c->dt = 0.002;
  c->pml_order = 4.0;  // DTmxw.cu:224
  c->pml_R = 1e-4;     // DTmxw.cu:225 (log(10000))
  c->pml_vmax = 1.0;
  c->nmat = 1;
  c->mat[0].eps_inf = 1.0;
  c->tfsf.amp = 1.0;
  c->tfsf.cSrc = 1.0;
  c->tfsf.Imp = 1.0;
  c->tfsf.sf_inside = 1;
  c->ntime = 1;
}

}  // extern "C"

// sigma u half-cells from the outer face of a layer of p cells.  POLY: the
// grading of DTmxw.cu:220-249 over the whole layer; REFERENCE: PMLgamma_func
// (DTmxw.cu:219-228) as setPMLcoeffs indexes it along x (DTmxw.cu:251-253),
// flat over the outer 3 half-cells, graded over N = 2p - 6, zero over the
// inner 3.
static double layer_sigma(int u, int p, double d, double m, double R, double vmax, int profile) {
  if (profile == DTM_PML_PROFILE_REFERENCE) {
    const double N = 2.0 * p - 6, a = std::max(u - 3, 0);
    if (a >= N) return 0.0;
    return vmax * log(1.0 / R) * (m + 1) / (N * d) * pow(1.0 - a / N, m);
  }
  const double smax = vmax * log(1.0 / R) * (m + 1) / (2.0 * p * d);
  return smax * pow((2.0 * p - u) / (2.0 * p), m);
}
// sigma at half-coordinate h of an axis of n cells with layers of plo / phi cells
static double pml_sigma(int h, int n, int plo, int phi, double d, double m, double R, double vmax, int profile) {
  if (plo > 0 && h < 2 * plo) return layer_sigma(h, plo, d, m, R, vmax, profile);
  if (phi > 0 && h > 2 * (n - phi)) return layer_sigma(2 * n - h, phi, d, m, R, vmax, profile);
  return 0.0;
}

static void ade_coeffs(const dtm_material& m, double dt, AdeMat& out) {
  // DispStruct::set (dispersion.cpp:41-82)
  double zm = 0, zp = 0, z0 = 0;
  double al[kMaxPoles], xi[kMaxPoles], zetp[kMaxPoles], zetm[kMaxPoles], zet[kMaxPoles];
  const int np = std::min(m.npoles, kMaxPoles);
  for (int i = 0; i < np; i++) {
    const LorenzDisp& L = m.pole[i];
    const double den = L.b1 * dt + 2;
    al[i] = (4 - 2 * L.b0 * dt * dt) / den;
    xi[i] = (L.b1 * dt - 2) / den;
    zetp[i] = (L.a0 * dt * dt + 2 * L.a1 * dt) / den;
    zetm[i] = (-L.a0 * dt * dt + 2 * L.a1 * dt) / den;
    zet[i] = -(4 * L.a1 * dt) / den;
    zm += zetm[i];
    zp += zetp[i];
    z0 += zet[i];
  }
  const double D = 2 * m.eps_inf + m.sigma * dt + zp;
  const double C1 = -zm / D, C2 = (2 * m.eps_inf - m.sigma * dt - z0) / D, C3 = 2 * dt / D;
  out.np = np;
  out.kE[0] = (float)C1;
  out.kE[1] = (float)C2;
  out.kE[2] = (float)(C3 / dt);
  for (int i = 0; i < np; i++) {
    out.kEJ[i][0] = (float)(C3 * 0.5 * (1 + al[i]));
    out.kEJ[i][1] = (float)(C3 * 0.5 * xi[i]);
    out.kJ[i][0] = (float)al[i];
    out.kJ[i][1] = (float)xi[i];
    out.kJ[i][2] = (float)(zetp[i] / dt);
    out.kJ[i][3] = (float)(zetm[i] / dt);
    out.kJ[i][4] = (float)(zet[i] / dt);
  }
}

// Magnetised-plasma current update matrices (trapezoidal rotation, no
// reference counterpart: the reference's Gyration() is disabled).
static void gyro_coeffs(const dtm_material& m, double dt, GyroMat& out) {
  const double w[3] = {m.gyro_wc[0], m.gyro_wc[1], m.gyro_wc[2]};
  const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double M[9], N[9];
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) {
      const double I = r == q ? 1.0 : 0.0;
      M[3 * r + q] = I * (1 + 0.5 * m.gyro_nu * dt) - 0.5 * dt * W[3 * r + q];
      N[3 * r + q] = I * (1 - 0.5 * m.gyro_nu * dt) + 0.5 * dt * W[3 * r + q];
    }
  const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
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
      out.A[3 * r + q] = (float)a;
      out.B[3 * r + q] = (float)(dt * m.gyro_wp * m.gyro_wp * Mi[3 * r + q]);
    }
}

static double wave_value(int waveform, double w, double kEnv, double t0, double tw, double t) {
  const double p = w * t;
  if (waveform == DTM_WAVE_RAISED_COS) {
    const double kp = p * kEnv;
    if (kp < 0 || kp > 2 * M_PI) return 0;
    return sin(p) * 0.5 * (1 - cos(kp));
  }
  const double u = (p / w - t0) / tw;
  if (waveform == DTM_WAVE_GAUSS) return exp(-u * u);
  return exp(-u * u) * sin(p);
}

// inSF bounds on half-coordinates: inside <=> h*d/2 > BoxM && h*d/2 < BoxP,
// evaluated with the same double expression as signal.hpp:154-157.
static void sf_bounds(double bm, double bp, double d, int nh, int& lo, int& hi) {
  lo = 1 << 30;
  hi = -(1 << 30);
  for (int h = -8; h <= nh + 8; h++) {
    const double x = h * 0.5 * d;
    if (x > bm && x < bp) {
      lo = std::min(lo, h);
      hi = std::max(hi, h);
    }
  }
}

// TF/SF geometry, incident-field constants and face bands of a config:
// host arithmetic only, so dtm_create validates the source before it
// touches the device.
static int tfsf_plan(const dtm_config& c, Tfsf& t, TfsfBands& B) {
  memset(&t, 0, sizeof(t));
  memset(&B, 0, sizeof(B));
  t.kind = c.tfsf.kind;
  if (t.kind == DTM_TFSF_NONE) return 0;
  if (t.kind != DTM_TFSF_DIPOLE && t.kind != DTM_TFSF_PLANE) return fail(DTM_ERR_ARG, "unknown tfsf kind");
  t.waveform = c.tfsf.waveform;
  t.sf_inside = c.tfsf.sf_inside;
  const int n[3] = {c.nx, c.ny, c.nz};
  const double d[3] = {c.dx, c.dy, c.dz};
  for (int a = 0; a < 3; a++) {
    sf_bounds(c.tfsf.box[2 * a], c.tfsf.box[2 * a + 1], d[a], 2 * n[a], t.hlo[a], t.hhi[a]);
    t.half_d[a] = 0.5 * d[a];
    if (t.hlo[a] > t.hhi[a]) {  // empty box: no face anywhere
      t.hlo[a] = 1 << 29;
      t.hhi[a] = -(1 << 29);
      t.clo[a] = 1;
      t.chi[a] = 0;
      continue;
    }
    // Cells whose stencil (half-offsets +-1, +-3 along a) can cross a face:
    // component half-coordinate h in [hlo-3, hlo+2] or [hhi-2, hhi+3], and
    // on the other axes h inside [hlo, hhi] (h = 2i or 2i+1).
    auto fdiv = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
    t.clo[a] = fdiv(t.hlo[a]);
    t.chi[a] = fdiv(t.hhi[a]);
    t.fl0[a] = fdiv(t.hlo[a] - 3);
    t.fl1[a] = fdiv(t.hlo[a] + 2);
    t.fh0[a] = fdiv(t.hhi[a] - 2);
    t.fh1[a] = fdiv(t.hhi[a] + 3);
  }
  t.w = c.tfsf.w;
  t.kEnv = c.tfsf.kEnv;
  t.cSrc = c.tfsf.cSrc;
  t.Imp = c.tfsf.Imp;
  t.t0 = c.tfsf.t0;
  t.tw = c.tfsf.tw;
  t.amp = c.tfsf.amp;
  t.st = sin(c.tfsf.theta);
  t.ct = cos(c.tfsf.theta);
  t.sph = sin(c.tfsf.phi);
  t.cph = cos(c.tfsf.phi);
  t.sps = sin(c.tfsf.psi);
  t.cps = cos(c.tfsf.psi);
  const double k = c.tfsf.w / c.tfsf.cSrc;
  t.kx = k * t.st * t.cph;
  t.ky = k * t.st * t.sph;
  t.kz = k * t.ct;
  for (int a = 0; a < 3; a++) {
    t.pos[a] = c.tfsf.pos[a];
    t.dip[a] = 0.5 * (c.tfsf.box[2 * a] + c.tfsf.box[2 * a + 1]);
    // The dipole field is singular at the box centre.  The corrections
    // evaluate it up to 2 half-cells inside a face, so the centre must lie
    // further in than that on every axis.
    const double hc = 2.0 * t.dip[a] / d[a];
    if (t.kind == DTM_TFSF_DIPOLE && t.hlo[a] <= t.hhi[a] && (hc - t.hlo[a] <= 2.5 || t.hhi[a] - hc <= 2.5))
      return fail(DTM_ERR_ARG, "dipole TF/SF box too small: its centre must lie more than 1.25 cells inside "
                               "every face (the incident field is singular there)");
  }
  // A plane wave whose k lies along one axis (the other components below
  // 1e-12 |k|, e.g. cos(pi/2) rounding) has a phase that depends on that
  // axis only: tabulate the waveform once per half-step.
  t.tab_axis = -1;
  if (t.kind == DTM_TFSF_PLANE) {
    const double ka[3] = {fabs(t.kx), fabs(t.ky), fabs(t.kz)};
    for (int a = 0; a < 3; a++) {
      const int b = (a + 1) % 3, e = (a + 2) % 3;
      if (ka[a] > 0 && ka[b] <= 1e-12 * ka[a] && ka[e] <= 1e-12 * ka[a]) t.tab_axis = a;
    }
    if (t.tab_axis >= 0) {
      t.tab_h0 = -8;
      t.tab_n = 2 * n[t.tab_axis] + 16;
      // incident component = P[c] * amp*f0(phase) (signal.hpp:28-49 with
      // E_phi = Imp cos(psi) s, E_theta = Imp sin(psi) s, H_theta = -cos(psi) s,
      // H_phi = sin(psi) s)
      t.P[0] = t.Imp * (-t.cps * t.sph + t.sps * t.ct * t.cph);
      t.P[1] = t.Imp * (t.cps * t.cph + t.sps * t.ct * t.sph);
      t.P[2] = -t.Imp * t.sps * t.st;
      t.P[3] = -t.sps * t.sph - t.cps * t.ct * t.cph;
      t.P[4] = t.sps * t.cph - t.cps * t.ct * t.sph;
      t.P[5] = t.cps * t.st;
    }
  }
  // face bands, clamped to the grid; lo/hi merged when they meet
  B.n = 0;
  auto clampr = [](int lo, int hi, int n, int& o0, int& o1) {
    o0 = std::max(lo, 0);
    o1 = std::min(hi + 1, n);
    return o1 > o0;
  };
  for (int a = 0; a < 3; a++) {
    if (t.hlo[a] > t.hhi[a]) continue;
    int ext0[3], ext1[3];
    bool ok = true;
    for (int b = 0; b < 3; b++) ok = ok && clampr(t.clo[b], t.chi[b], n[b], ext0[b], ext1[b]);
    if (!ok) continue;
    int bands[2][2];
    int nb = 0;
    if (t.fh0[a] <= t.fl1[a] + 1) {
      if (clampr(t.fl0[a], t.fh1[a], n[a], bands[0][0], bands[0][1])) nb = 1;
    } else {
      if (clampr(t.fl0[a], t.fl1[a], n[a], bands[nb][0], bands[nb][1])) nb++;
      if (clampr(t.fh0[a], t.fh1[a], n[a], bands[nb][0], bands[nb][1])) nb++;
    }
    for (int q = 0; q < nb; q++) {
      int r0[3] = {ext0[0], ext0[1], ext0[2]}, r1[3] = {ext1[0], ext1[1], ext1[2]};
      r0[a] = bands[q][0];
      r1[a] = bands[q][1];
      B.r[B.n] = Region{r0[0], r1[0], r0[1], r1[1], r0[2], r1[2]};
      B.axis[B.n] = a;
      B.n++;
    }
  }
  // the correction kernel indexes band cells with 32-bit thread ids, each
  // band padded to whole wavefronts
  long long nthr = 0;
  for (int r = 0; r < B.n; r++) {
    const Region& R = B.r[r];
    nthr += ((long long)(R.i1 - R.i0) * (R.j1 - R.j0) * (R.k1 - R.k0) + 63) / 64 * 64;
  }
  if (nthr >= (1LL << 31))
    return fail(DTM_ERR_LIMIT, "TF/SF face bands hold 2^31 cells or more (the correction kernel's index range)");
  return 0;
}

static int setup_tfsf(dtm_solver* s) {
  if (int rc = tfsf_plan(s->cfg, s->tf, s->bands)) return rc;
  s->tf.ptab = nullptr;
  if (s->tf.kind == DTM_TFSF_DIPOLE) {
    // every incident value the corrections read lies at half-coordinates
    // [hlo - 3, hhi + 3] on each axis (a band cell's neighbour across a face
    // is at most 3 half-cells from it, and the other axes lie in the box)
    long long tot = 6;
    for (int a = 0; a < 3; a++) {
      const long long n = (long long)s->tf.hhi[a] - s->tf.hlo[a] + 7;
      tot = (n <= 0 || n > kMaxPointTable) ? kMaxPointTable + 1 : std::min(tot * n, kMaxPointTable + 1);
      s->tf.ptab_h0[a] = s->tf.hlo[a] - 3;
      s->tf.ptab_n[a] = (int)std::max(0LL, std::min(n, kMaxPointTable));
    }
    if (tot <= kMaxPointTable)
      if (int rc = dalloc(s, &s->tf.ptab, sizeof(double) * (size_t)tot)) return rc;
  }
  if (s->tf.kind != DTM_TFSF_NONE && s->tf.tab_axis >= 0) {
    if (int rc = dalloc(s, &s->tab_base, sizeof(double) * 2 * kTableSteps * s->tf.tab_n)) return rc;
    // per-plane factors of a wave along x, for the sweeps (and the fused step) that fold it in
    if (s->tf.tab_axis == 0)
      if (int rc = dalloc(s, &s->fold_base, sizeof(float) * 2 * kTableSteps * fold_floats(s))) return rc;
  }
  return 0;
}

// Tfsf of one TF/SF launch: for a tabulated plane wave, points the table
// at step n, half (0: H half-step, 1: E half-step), filling the next batch
// of tables first when step n is not in the current one.
static Tfsf tfsf_for(dtm_solver* s, long long n, int half) {
  Tfsf t = s->tf;
  if (t.tab_axis < 0) return t;
  if (s->tab_n0 < 0 || n < s->tab_n0 || n >= s->tab_n0 + kTableSteps) {
    s->tab_n0 = n;
    launch_tfsf_tables(s->tf, s->tab_base, n, kTableSteps, s->cfg.dt, s->stream);
    if (s->fold_base) launch_tfsf_fold(s->tf, s->tab_base, s->fold_base, kTableSteps, s->g, s->stream);
  }
  t.table = s->tab_base + ((n - s->tab_n0) * 2 + half) * (long long)t.tab_n;
  return t;
}

// The fold factor table of step n, half (after tfsf_for filled its batch).
static const float* fold_for(const dtm_solver* s, long long n, int half) {
  return s->fold_base + ((n - s->tab_n0) * 2 + half) * (long long)fold_floats(s);
}

static bool tfsf_gate(const dtm_solver* s, long long n) {
  const dtm_config& c = s->cfg;
  if (c.tfsf.kind == DTM_TFSF_NONE) return false;
  const int nt = c.ntime > 0 ? c.ntime : 1;
  const long long istep = n / nt;
  return (double)((istep + 1) * nt) * c.dt < c.tfsf.tStop;  // window.hpp:125
}

static int setup_cpml(dtm_solver* s) {
  const dtm_config& c = s->cfg;
  Cpml& p = s->pm;
  const int n[3] = {c.nx, c.ny, c.nz};
  const double d[3] = {c.dx, c.dy, c.dz};
  p.on = 0;
  for (int a = 0; a < 3; a++) {
    p.plo[a] = c.npml[2 * a];
    p.phi[a] = c.npml[2 * a + 1];
    if (p.plo[a] > 0 || p.phi[a] > 0) p.on = 1;
  }
  p.split = (p.on && c.pml_kind == DTM_PML_SPLIT) ? 1 : 0;
  for (int a = 0; a < 3; a++) {
    std::vector<float> bE(n[a]), cE(n[a]), bH(n[a]), cH(n[a]);
    for (int v = 0; v < n[a]; v++) {
      const double se =
          pml_sigma(2 * v + 1, n[a], p.plo[a], p.phi[a], d[a], c.pml_order, c.pml_R, c.pml_vmax, c.pml_profile);
      const double sh =
          pml_sigma(2 * v, n[a], p.plo[a], p.phi[a], d[a], c.pml_order, c.pml_R, c.pml_vmax, c.pml_profile);
      if (p.split) {  // setPMLcoeffs (DTmxw.cu:250-265)
        const double k2e = 1.0 / (1.0 + 0.5 * c.dt * se), k2h = 1.0 / (1.0 + 0.5 * c.dt * sh);
        bE[v] = (float)(2.0 * k2e - 1.0);
        cE[v] = (float)k2e;
        bH[v] = (float)(2.0 * k2h - 1.0);
        cH[v] = (float)k2h;
        continue;
      }
      const double be = exp(-se * c.dt), bh = exp(-sh * c.dt);
      bE[v] = (float)be;
      cE[v] = (float)(be - 1.0);
      bH[v] = (float)bh;
      cH[v] = (float)(bh - 1.0);
    }
    float* t[4];
    for (int q = 0; q < 4; q++)
      if (int rc = dalloc(s, &t[q], sizeof(float) * n[a])) return rc;
    HIPCHK(hipMemcpy(t[0], bE.data(), sizeof(float) * n[a], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t[1], cE.data(), sizeof(float) * n[a], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t[2], bH.data(), sizeof(float) * n[a], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t[3], cH.data(), sizeof(float) * n[a], hipMemcpyHostToDevice));
    p.bE[a] = t[0];
    p.cE[a] = t[1];
    p.bH[a] = t[2];
    p.cH[a] = t[3];
    const int P = p.plo[a] + p.phi[a];
    size_t cnt = 0;
    if (P > 0 && !p.split)
      cnt = (size_t)P * (a == 0 ? (size_t)c.ny * c.nz : a == 1 ? (size_t)c.nx * c.nz : (size_t)c.nx * c.ny);
    if (cnt >= (1ull << 32)) return fail(DTM_ERR_LIMIT, "a CPML slab holds 2^32 cells or more (32-bit psi indices)");
    for (int sl = 0; sl < 2; sl++) {
      if (int rc = sdalloc(s, &p.psiE[a][sl], sizeof(float) * cnt)) return rc;
      if (int rc = sdalloc(s, &p.psiH[a][sl], sizeof(float) * cnt)) return rc;
    }
  }
  if (p.split) {
    const size_t cells = (size_t)c.nx * c.ny * c.nz + kPad;
    for (int q = 0; q < 3; q++) {
      if (int rc = sdalloc(s, &p.p2E[q], sizeof(float) * cells)) return rc;
      if (int rc = sdalloc(s, &p.p2H[q], sizeof(float) * cells)) return rc;
    }
  }
  s->pmn = p;
  if (s->torre) {
    for (int a = 0; a < 3; a++) {
      const int P = p.plo[a] + p.phi[a];
      size_t cnt = 0;
      if (P > 0) cnt = (size_t)P * (a == 0 ? (size_t)c.ny * c.nz : a == 1 ? (size_t)c.nx * c.nz : (size_t)c.nx * c.ny);
      for (int sl = 0; sl < 2; sl++) {
        if (int rc = sdalloc(s, &s->pmn.psiE[a][sl], sizeof(float) * cnt)) return rc;
        if (int rc = sdalloc(s, &s->pmn.psiH[a][sl], sizeof(float) * cnt)) return rc;
      }
    }
  }
  return 0;
}

// Launch layout.  The fused step and the sweeps run the whole grid in one
// launch of the variant that applies every boundary rule (its x, y and z
// CPML branches are plane-, row- and column-uniform): at 512x256x256 one
// launch beat an interior launch plus a y-strip launch by 7-9% (FP32),
// 4% (FP16 H) and 4.5% (FP16 E), because the strips' few tile columns
// forced short x segments.
static void setup_regions(dtm_solver* s) {
  const dtm_config& c = s->cfg;
  s->whole = Region{0, c.nx, 0, c.ny, 0, c.nz};
}

static void field_structs(dtm_solver* s) {
  for (int c = 0; c < 3; c++) {
    s->f32.E[c] = static_cast<float*>(s->E[c]);
    s->f32.H[c] = static_cast<float*>(s->H[c]);
    s->f16.E[c] = static_cast<__half*>(s->E[c]);
    s->f16.H[c] = static_cast<__half*>(s->H[c]);
    s->f32n.E[c] = static_cast<float*>(s->E2[c]);
    s->f32n.H[c] = static_cast<float*>(s->H2[c]);
    s->f16n.E[c] = static_cast<__half*>(s->E2[c]);
    s->f16n.H[c] = static_cast<__half*>(s->H2[c]);
  }
}

static const void* fields_ptr(const dtm_solver* s) {
  return s->store == 0 ? static_cast<const void*>(&s->f32) : static_cast<const void*>(&s->f16);
}
static const void* fields_next_ptr(const dtm_solver* s) {
  return s->store == 0 ? static_cast<const void*>(&s->f32n) : static_cast<const void*>(&s->f16n);
}

// After a fused step the new buffers become current.
static void swap_buffers(dtm_solver* s) {
  for (int c = 0; c < 3; c++) {
    std::swap(s->E[c], s->E2[c]);
    std::swap(s->H[c], s->H2[c]);
  }
  field_structs(s);
  for (int a = 0; a < 3; a++)
    for (int sl = 0; sl < 2; sl++) {
      std::swap(s->pm.psiE[a][sl], s->pmn.psiE[a][sl]);
      std::swap(s->pm.psiH[a][sl], s->pmn.psiH[a][sl]);
    }
}

static int build_classes_into(dtm_solver* s, const uint8_t* const m[3]);
// class of cell q (an empty class array: every cell is class 0)
static int cls_of(const dtm_solver* s, size_t q) { return s->cls_host.empty() ? 0 : s->cls_host[q]; }

// Build the class table from per-component material ids.  A volume that is
// rejected leaves the solver's current materials in place.
static int build_classes(dtm_solver* s, const uint8_t* const m[3]) {
  auto classes = s->classes;
  auto class_mats = s->class_mats;
  auto cls_host = s->cls_host;
  const int rc = build_classes_into(s, m);
  if (rc) {
    s->classes.swap(classes);
    s->class_mats.swap(class_mats);
    s->cls_host.swap(cls_host);
  }
  return rc;
}

static int build_classes_into(dtm_solver* s, const uint8_t* const m[3]) {
  const dtm_config& c = s->cfg;
  const size_t n = (size_t)c.nx * c.ny * c.nz;
  std::unordered_map<uint32_t, int> ids;
  s->classes.clear();
  s->class_mats.clear();
  s->cls_host.clear();
  auto entry_for = [&](uint8_t a, uint8_t b, uint8_t d) -> int {
    const uint32_t key = a | (b << 8) | (d << 16);
    auto it = ids.find(key);
    if (it != ids.end()) return it->second;
    const int id = (int)s->classes.size();
    if (id >= kMaxClasses) return -1;
    ids[key] = id;
    ClassEntry ce{};
    const uint8_t mm[3] = {a, b, d};
    for (int q = 0; q < 3; q++) {
      const dtm_material& M = c.mat[mm[q]];
      ce.inv_eps[q] = (float)(1.0 / (M.eps_inf > 0 ? M.eps_inf : 1.0));
      ce.mat[q] = mm[q];
      if (M.npoles > 0 || M.sigma != 0) ce.special |= (uint8_t)(1 << q);
    }
    if (c.mat[a].gyro) ce.special |= 8;
    s->classes.push_back(ce);
    s->class_mats.push_back({a, b, d});
    return id;
  };
  entry_for(0, 0, 0);
  int bb[6] = {c.nx, -1, c.ny, -1, c.nz, -1};
  bool any_special = false, any_gyro = false;
  // no volumes and a plain material 0: every cell is class 0, and no
  // per-cell class array exists (cls_of)
  const bool uniform = !m[0] && !m[1] && !m[2] && !s->classes[0].special;
  if (!uniform) s->cls_host.assign(n, 0);
  int cx0 = c.nx, cx1 = 0;  // x range of cells whose class is not 0
  for (size_t q = 0; q < (uniform ? 0 : n); q++) {
    const uint8_t a = m[0] ? m[0][q] : 0, b = m[1] ? m[1][q] : 0, d = m[2] ? m[2][q] : 0;
    if (a >= c.nmat || b >= c.nmat || d >= c.nmat) return fail(DTM_ERR_ARG, "material id >= nmat");
    const int id = entry_for(a, b, d);
    if (id < 0) return fail(DTM_ERR_LIMIT, "more than 256 distinct (Ex,Ey,Ez) material classes");
    s->cls_host[q] = (uint8_t)id;
    if (id) {
      const int i = (int)(q / ((size_t)c.ny * c.nz));
      cx0 = std::min(cx0, i);
      cx1 = std::max(cx1, i + 1);
    }
    if (s->classes[id].special) {
      any_special = true;
      if (s->classes[id].special & 8) any_gyro = true;
      const int i = (int)(q / ((size_t)c.ny * c.nz)), j = (int)((q / c.nz) % c.ny), k = (int)(q % c.nz);
      const int v[3] = {i, j, k}, nn[3] = {c.nx, c.ny, c.nz};
      bool in_layer = false;
      for (int a = 0; a < 3; a++)
        if (v[a] < c.npml[2 * a] || v[a] >= nn[a] - c.npml[2 * a + 1]) in_layer = true;
      if (in_layer && s->pm.split)
        return fail(DTM_ERR_ARG, "the split-field layer does not cover dispersive or gyrotropic cells");
      // A Drude/Lorentz cell inside a CPML slab makes the CPML-ADE update grow
      // without bound (a silver half-space in the layer: the oracle and the
      // GPU both diverge after ~2000 steps); the reference keeps its layer out
      // of the metal too (non-absorbing low-z face, DTmxw.cu:262).
      if (in_layer) {
        const auto& cm = s->class_mats[id];
        for (int q3 = 0; q3 < 3; q3++)
          if (c.mat[cm[q3]].npoles > 0)
            return fail(DTM_ERR_ARG, "a dispersive (Drude/Lorentz) cell lies in a CPML slab: leave that face "
                                     "without a layer (npml = 0) or keep the medium out of the layer");
      }
      bb[0] = std::min(bb[0], i);
      bb[1] = std::max(bb[1], i);
      bb[2] = std::min(bb[2], j);
      bb[3] = std::max(bb[3], j);
      bb[4] = std::min(bb[4], k);
      bb[5] = std::max(bb[5], k);
    }
  }
  if (s->mt.d_on && any_special) {  // a design box set earlier must stay on plain cells
    const Materials& mt = s->mt;
    for (int i = mt.d_i0; i < mt.d_i0 + mt.d_ni; i++)
      for (int j = mt.d_j0; j < mt.d_j0 + mt.d_nj; j++)
        for (int k = mt.d_k0; k < mt.d_k0 + mt.d_nk; k++)
          if (s->classes[cls_of(s, ((size_t)i * c.ny + j) * c.nz + k)].special)
            return fail(DTM_ERR_ARG, "the design box overlaps a dispersive or gyrotropic cell");
  }
  // soft sources must sit in plain cells
  for (int q = 0; q < c.nsoft; q++) {
    const dtm_soft_source& so = c.soft[q];
    const size_t idx = ((size_t)so.i * c.ny + so.j) * c.nz + so.k;
    if (s->classes[cls_of(s, idx)].special)
      return fail(DTM_ERR_ARG, "soft source placed in a dispersive/gyrotropic cell");
  }
  const int ncls = (int)s->classes.size();
  if (!s->table_dev)
    if (int rc = dalloc(s, &s->table_dev, sizeof(ClassEntry) * kMaxClasses)) return rc;
  HIPCHK(hipMemcpy(s->table_dev, s->classes.data(), sizeof(ClassEntry) * ncls, hipMemcpyHostToDevice));
  if (ncls > 1) {
    if (!s->cls_dev)
      if (int rc = dalloc(s, &s->cls_dev, n + kPad)) return rc;
    HIPCHK(hipMemcpy(s->cls_dev, s->cls_host.data(), n, hipMemcpyHostToDevice));
  }
  s->mt.cls = ncls > 1 ? s->cls_dev : nullptr;
  s->mt.cx0 = ncls > 1 ? cx0 : 0;
  s->mt.cx1 = ncls > 1 ? cx1 : 0;
  s->mt.table = s->table_dev;
  s->mt.nclass = ncls;
  // dispersive / gyrotropic state over the bounding box of special cells
  Ade& A = s->ade;
  A.on = any_special ? 1 : 0;
  A.gyro_on = any_gyro ? 1 : 0;
  if (any_special) {
    A.i0 = bb[0];
    A.j0 = bb[2];
    A.k0 = bb[4];
    A.ni = bb[1] - bb[0] + 1;
    A.nj = bb[3] - bb[2] + 1;
    A.nk = bb[5] - bb[4] + 1;
    A.bsy = A.nk;
    A.bsx = (long long)A.nj * A.nk;
    const size_t nb = (size_t)A.ni * A.nj * A.nk;
    int maxp = 0;
    std::vector<AdeMat> am(DTM_MAX_MATS);
    std::vector<GyroMat> gm(DTM_MAX_MATS);
    for (int mm = 0; mm < c.nmat; mm++) {
      ade_coeffs(c.mat[mm], c.dt, am[mm]);
      maxp = std::max(maxp, am[mm].np);
      if (c.mat[mm].gyro) gyro_coeffs(c.mat[mm], c.dt, gm[mm]);
    }
    for (int q = 0; q < 3; q++) {
      if (int rc = dalloc(s, &A.curl[q], sizeof(float) * nb)) return rc;
      if (int rc = sdalloc(s, &A.Em[q], sizeof(float) * nb)) return rc;
      for (int p = 0; p < maxp; p++) {
        if (int rc = sdalloc(s, &A.Jp[q][p], sizeof(float) * nb)) return rc;
        if (int rc = sdalloc(s, &A.Jm[q][p], sizeof(float) * nb)) return rc;
      }
      if (any_gyro)
        if (int rc = sdalloc(s, &A.Jg[q], sizeof(float) * nb)) return rc;
    }
    if (int rc = dalloc(s, &s->ademats_dev, sizeof(AdeMat) * DTM_MAX_MATS)) return rc;
    if (int rc = dalloc(s, &s->gyromats_dev, sizeof(GyroMat) * DTM_MAX_MATS)) return rc;
    HIPCHK(hipMemcpy(s->ademats_dev, am.data(), sizeof(AdeMat) * DTM_MAX_MATS, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->gyromats_dev, gm.data(), sizeof(GyroMat) * DTM_MAX_MATS, hipMemcpyHostToDevice));
    A.mats = s->ademats_dev;
    A.gmats = s->gyromats_dev;
  }
  return 0;
}

extern "C" {

int dtm_create(const dtm_config* cfg, dtm_solver** out) {
  if (!cfg || !out) return fail(DTM_ERR_ARG, "null argument");
  *out = nullptr;
  const dtm_config& c = *cfg;
  if (c.nx < 1 || c.ny < 1 || c.nz < 1) return fail(DTM_ERR_ARG, "grid sizes must be positive");
  // 2^32 cells and more: the sweeps index cells with 64 bits (their wide
  // variants), which the split-field layer, the FP32 zone and the fused step
  // do not have
  const bool wide = (unsigned long long)c.nx * c.ny * c.nz > 0xffffffffULL - kPad;
  if (wide && c.pml_kind == DTM_PML_SPLIT)
    return fail(DTM_ERR_LIMIT, "the split-field layer is limited to grids below 2^32 cells");
  if ((long long)c.nx * c.ny * c.nz > (1LL << 34)) return fail(DTM_ERR_LIMIT, "grid too large");
  if (!(c.dx > 0 && c.dy > 0 && c.dz > 0 && c.dt > 0)) return fail(DTM_ERR_ARG, "steps must be positive");
  const double cour = c.dt * sqrt(1 / (c.dx * c.dx) + 1 / (c.dy * c.dy) + 1 / (c.dz * c.dz));
  if (cour > 6. / 7.) return fail(DTM_ERR_COURANT, "Courant condition is not satisfied (DTmxw.cu:332)");
  if (c.nmat < 1 || c.nmat > DTM_MAX_MATS) return fail(DTM_ERR_ARG, "nmat out of range");
  if (c.nsoft < 0 || c.nsoft > DTM_MAX_SOFT) return fail(DTM_ERR_ARG, "nsoft out of range");
  if (c.storage != DTM_STORE_FP32 && c.storage != DTM_STORE_FP16) return fail(DTM_ERR_ARG, "bad storage");
  if (c.sweep != DTM_SWEEP_TWOPASS && c.sweep != DTM_SWEEP_FUSED) return fail(DTM_ERR_ARG, "bad sweep");
  if (c.pml_kind != DTM_PML_CPML && c.pml_kind != DTM_PML_SPLIT) return fail(DTM_ERR_ARG, "bad pml_kind");
  if (c.pml_profile != DTM_PML_PROFILE_POLY && c.pml_profile != DTM_PML_PROFILE_REFERENCE)
    return fail(DTM_ERR_ARG, "bad pml_profile");
  if (c.pml_profile == DTM_PML_PROFILE_REFERENCE)
    for (int q = 0; q < 6; q++)
      if (c.npml[q] > 0 && c.npml[q] < 4)
        return fail(DTM_ERR_ARG, "the reference sigma profile needs layers of at least 4 cells");
  if (c.pml_kind == DTM_PML_SPLIT && c.sweep == DTM_SWEEP_FUSED)
    return fail(DTM_ERR_ARG, "the split-field layer runs with the two-pass sweeps");
  if (c.sweep == DTM_SWEEP_FUSED &&
      6ull * ((unsigned long long)c.nx * c.ny * c.nz + kPad) * (c.storage ? 2 : 4) >= (1ull << 32))
    return fail(DTM_ERR_LIMIT, "the fused step addresses a field set (six arrays) with 32-bit byte offsets: "
                               "at most 4 GiB, use the two-pass sweeps for larger grids");
  const int n3[3] = {c.nx, c.ny, c.nz};
  for (int a = 0; a < 3; a++) {
    if (c.npml[2 * a] < 0 || c.npml[2 * a + 1] < 0 || c.npml[2 * a] + c.npml[2 * a + 1] > n3[a])
      return fail(DTM_ERR_ARG, "CPML thicker than the grid");
    if (c.bc[a] == DTM_BC_PERIODIC && (c.npml[2 * a] || c.npml[2 * a + 1]))
      return fail(DTM_ERR_ARG, "CPML on a periodic axis");
    if (c.bc[a] == DTM_BC_PERIODIC && n3[a] < 4) return fail(DTM_ERR_ARG, "periodic axis needs >= 4 cells");
  }
  for (int m = 0; m < c.nmat; m++) {
    if (!(c.mat[m].eps_inf > 0)) return fail(DTM_ERR_ARG, "eps_inf must be positive");
    if (c.mat[m].npoles < 0 || c.mat[m].npoles > DTM_MAX_POLES) return fail(DTM_ERR_ARG, "npoles out of range");
  }
  for (int q = 0; q < c.nsoft; q++) {
    const dtm_soft_source& so = c.soft[q];
    if (so.comp < 0 || so.comp > 2 || so.i < 0 || so.i >= c.nx || so.j < 0 || so.j >= c.ny || so.k < 0 ||
        so.k >= c.nz)
      return fail(DTM_ERR_ARG, "soft source outside the grid");
  }
  {  // the source is checked before any device call
    Tfsf t;
    TfsfBands b;
    if (int rc = tfsf_plan(c, t, b)) return rc;
  }
  dtm_solver* s = new dtm_solver();
  s->cfg = c;
  s->store = c.storage;
  s->torre = c.sweep == DTM_SWEEP_FUSED;
  hipError_t e = hipGetDevice(&s->dev);
  if (e != hipSuccess) {
    delete s;
    return fail(DTM_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
  }
  e = hipStreamCreate(&s->stream);
  if (e != hipSuccess) {
    delete s;
    return fail(DTM_ERR_DEVICE, hipGetErrorString(e));
  }
  if (static_scratch_init(s->dev)) {  // device address of dtm_energy_box_noalloc's static scratch
    (void)hipStreamDestroy(s->stream);
    delete s;
    return fail(DTM_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed");
  }
  Grid& g = s->g;
  g.nx = c.nx;
  g.ny = c.ny;
  g.nz = c.nz;
  g.sy = c.nz;
  g.sx = (long long)c.ny * c.nz;
  for (int a = 0; a < 3; a++) g.per[a] = c.bc[a] == DTM_BC_PERIODIC;
  g.rd[0] = (float)(1.0 / (24.0 * c.dx));
  g.rd[1] = (float)(1.0 / (24.0 * c.dy));
  g.rd[2] = (float)(1.0 / (24.0 * c.dz));
  g.dt = (float)c.dt;
  for (int a = 0; a < 3; a++) {
    const int se = s->store ? c.scale_e_log2[a] : 0, sh = s->store ? c.scale_h_log2[a] : 0;
    s->sc.se[a] = ldexpf(1.f, se);
    s->sc.ie[a] = ldexpf(1.f, -se);
    s->sc.sh[a] = ldexpf(1.f, sh);
    s->sc.ih[a] = ldexpf(1.f, -sh);
  }
  {
    void* z = nullptr;
    if (int rc = dalloc(s, &z, 256)) {
      dtm_destroy(s);
      return rc;
    }
    g.zero = z;
  }
  g.zidx = (unsigned)((size_t)c.nx * c.ny * c.nz);  // wide grids use zidxw
  g.zidxw = (long long)((size_t)c.nx * c.ny * c.nz);
  const size_t n = (size_t)c.nx * c.ny * c.nz;
  const size_t esz = s->store ? 2 : 4;
  // the six components of a field set share one allocation (Ex Ey Ez Hx Hy
  // Hz, each n + kPad elements), so the fused step addresses all of them from
  // one base with 32-bit byte offsets
  auto field_set = [&](void* e[3], void* h[3]) -> int {
    const size_t per = esz * (n + kPad);
    char* b = nullptr;
    if (int rc2 = dalloc(s, &b, 6 * per)) return rc2;
    for (int q = 0; q < 3; q++) {
      e[q] = b + q * per;
      h[q] = b + (3 + q) * per;
    }
    for (int q = 0; q < 3; q++) s->state.emplace_back(e[q], per);
    for (int q = 0; q < 3; q++) s->state.emplace_back(h[q], per);
    return 0;
  };
  int rc = field_set(s->E, s->H);
  if (!rc && s->torre) rc = field_set(s->E2, s->H2);
  if (!rc) rc = setup_cpml(s);
  if (!rc) rc = setup_tfsf(s);
  if (!rc) {
    const uint8_t* none[3] = {nullptr, nullptr, nullptr};
    rc = build_classes(s, none);
  }
  if (!rc) rc = dalloc(s, &s->energy_dev, sizeof(double));
  if (rc) {
    dtm_destroy(s);
    return rc;
  }
  field_structs(s);
  setup_regions(s);
  if (hipStreamSynchronize(s->stream) != hipSuccess) {
    dtm_destroy(s);
    return fail(DTM_ERR_DEVICE, "stream sync failed after allocation");
  }
  *out = s;
  return 0;
}

void dtm_destroy(dtm_solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& e : s->ev_pool) {
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
  }
  if (s->side.st) (void)hipStreamSynchronize(s->side.st);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->side.fork) (void)hipEventDestroy(s->side.fork);
  if (s->side.join) (void)hipEventDestroy(s->side.join);
  if (s->side.st) (void)hipStreamDestroy(s->side.st);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int dtm_set_materials(dtm_solver* s, const uint8_t* mx, const uint8_t* my, const uint8_t* mz) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  if (s->started) return fail(DTM_ERR_STATE, "materials must be set before the first step");
  const uint8_t* m[3] = {mx, my, mz};
  return build_classes(s, m);
}

int dtm_set_design_eps(dtm_solver* s, int i0, int j0, int k0, int ni, int nj, int nk, const double* eps) {
  if (!s || !eps) return fail(DTM_ERR_ARG, "null argument");
  const dtm_config& c = s->cfg;
  if (i0 < 0 || j0 < 0 || k0 < 0 || ni < 1 || nj < 1 || nk < 1 || i0 + ni > c.nx || j0 + nj > c.ny || k0 + nk > c.nz)
    return fail(DTM_ERR_ARG, "design box outside the grid");
  const size_t nb = (size_t)ni * nj * nk;
  if (s->mt.d_on && (s->mt.d_ni != ni || s->mt.d_nj != nj || s->mt.d_nk != nk))
    return fail(DTM_ERR_ARG, "design box shape cannot change");
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++) {
        const size_t q = ((size_t)(i0 + i) * c.ny + (j0 + j)) * c.nz + (k0 + k);
        const ClassEntry& ce = s->classes[cls_of(s, q)];
        if (ce.special) return fail(DTM_ERR_ARG, "design box overlaps a dispersive cell");
      }
  s->design_host.resize(nb);
  for (size_t q = 0; q < nb; q++) {
    if (!(eps[q] > 0)) return fail(DTM_ERR_ARG, "design eps must be positive");
    s->design_host[q] = (float)(1.0 / eps[q]);
  }
  if (!s->design_dev)
    if (int rc = dalloc(s, &s->design_dev, sizeof(float) * nb)) return rc;
  HIPCHK(hipMemcpyAsync(s->design_dev, s->design_host.data(), sizeof(float) * nb, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->mt.d_on = 1;
  s->mt.d_i0 = i0;
  s->mt.d_j0 = j0;
  s->mt.d_k0 = k0;
  s->mt.d_ni = ni;
  s->mt.d_nj = nj;
  s->mt.d_nk = nk;
  s->mt.d_inv_eps = s->design_dev;
  return 0;
}

static int add_probe_desc(dtm_solver* s, const ProbeDesc& d) {
  if (s->started) return fail(DTM_ERR_STATE, "probes must be added before the first step");
  if ((int)s->probes.size() >= kMaxProbes) return fail(DTM_ERR_LIMIT, "too many probes");
  s->probes.push_back(d);
  return (int)s->probes.size() - 1;
}

int dtm_add_sensor(dtm_solver* s, const char* fld, int x, int y, int z) {
  if (!s || !fld) return fail(DTM_ERR_ARG, "null argument");
  static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  int comp = -1;
  for (int q = 0; q < 6; q++)
    if (strcmp(fld, names[q]) == 0) comp = q;
  if (comp < 0) return fail(DTM_ERR_ARG, "sensor field must be one of Ex Ey Ez Hx Hy Hz");
  const dtm_config& c = s->cfg;
  if (x < 0 || x >= c.nx || y < 0 || y >= c.ny || z < 0 || z >= c.nz) return fail(DTM_ERR_ARG, "sensor outside grid");
  return add_probe_desc(s, ProbeDesc{1, comp, x, y, z});
}

int dtm_add_probe(dtm_solver* s, int comp, int i, int j, int k) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  const dtm_config& c = s->cfg;
  if (comp < 0 || comp > 5 || i < 0 || i >= c.nx || j < 0 || j >= c.ny || k < 0 || k >= c.nz)
    return fail(DTM_ERR_ARG, "probe outside grid");
  return add_probe_desc(s, ProbeDesc{0, comp, i, j, k});
}

int dtm_set_dft(dtm_solver* s, int nfreq, const double* omegas) {
  if (!s || nfreq < 0 || (nfreq > 0 && !omegas)) return fail(DTM_ERR_ARG, "bad dft arguments");
  if (s->started) return fail(DTM_ERR_STATE, "DFT must be configured before the first step");
  if (nfreq > 4096) return fail(DTM_ERR_LIMIT, "too many DFT frequencies");
  s->nfreq = nfreq;
  if (nfreq == 0) return 0;
  if (int rc = dalloc(s, &s->omega_dev, sizeof(double) * nfreq)) return rc;
  HIPCHK(hipMemcpy(s->omega_dev, omegas, sizeof(double) * nfreq, hipMemcpyHostToDevice));
  return 0;
}

static int start(dtm_solver* s) {
  if (s->started) return 0;
  const int np = (int)s->probes.size();
  const int nf = (int)s->flux_faces.size();
  if (np > 0 || nf > 0) s->series_cap = std::max(16, std::min(4096, (1 << 22) / std::max(np, nf)));
  if (np > 0) {
    if (int rc = dalloc(s, &s->probes_dev, sizeof(ProbeDesc) * np)) return rc;
    HIPCHK(hipMemcpy(s->probes_dev, s->probes.data(), sizeof(ProbeDesc) * np, hipMemcpyHostToDevice));
    if (int rc = dalloc(s, &s->series_dev, sizeof(float) * (size_t)s->series_cap * np)) return rc;
    if (s->nfreq > 0)
      if (int rc = dalloc(s, &s->dft_dev, sizeof(double) * 2 * (size_t)np * s->nfreq)) return rc;
  }
  if (nf > 0) {
    FluxSet& fs = s->fs;
    fs.nface = nf;
    for (int q = 0; q < nf; q++) fs.face[q] = s->flux_faces[q];
    if (int rc = dalloc(s, &fs.snap, sizeof(float) * (size_t)s->flux_snap_n)) return rc;
    if (int rc = dalloc(s, &fs.series, sizeof(double) * (size_t)s->series_cap * nf)) return rc;
    if (int rc = dalloc(s, &fs.total, sizeof(double) * nf)) return rc;
  }
  s->started = true;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

static int flush_series(dtm_solver* s) {
  const int np = (int)s->probes.size();
  const int nf = s->fs.nface;
  if ((np == 0 && nf == 0) || s->series_row == 0) return 0;
  if (np > 0) {
    const size_t old = s->series_host.size();
    s->series_host.resize(old + (size_t)s->series_row * np);
    HIPCHK(hipMemcpyAsync(s->series_host.data() + old, s->series_dev, sizeof(float) * (size_t)s->series_row * np,
                          hipMemcpyDeviceToHost, s->stream));
  }
  if (nf > 0) {
    const size_t old = s->flux_host.size();
    s->flux_host.resize(old + (size_t)s->series_row * nf);
    HIPCHK(hipMemcpyAsync(s->flux_host.data() + old, s->fs.series, sizeof(double) * (size_t)s->series_row * nf,
                          hipMemcpyDeviceToHost, s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  s->series_row = 0;
  return 0;
}


static void ev_begin(dtm_solver* s, int group) {
  if (!s->timing) return;
  if (s->ev_used == s->ev_pool.size()) {
    std::array<hipEvent_t, 2> e{};
    (void)hipEventCreate(&e[0]);
    (void)hipEventCreate(&e[1]);
    s->ev_pool.push_back(e);
    s->ev_group.push_back(0);
  }
  s->ev_group[s->ev_used] = group;
  (void)hipEventRecord(s->ev_pool[s->ev_used][0], s->stream);
}
static void ev_end(dtm_solver* s) {
  if (!s->timing) return;
  (void)hipEventRecord(s->ev_pool[s->ev_used][1], s->stream);
  s->ev_used++;
}

// Device copy of the fused step's pointer table for the current buffers.
// The two slots match the two buffer parities, so uploads happen only when a
// pointer changes (materials, design box); the stream is drained first, as
// a queued step may still read the slot.
static const TorreCold* torre_cold(dtm_solver* s) {
  const TorreCold h = make_torre_cold(fields_ptr(s), fields_next_ptr(s), s->store, s->pm, s->pmn, s->mt, s->ade);
  for (int k = 0; k < 2; k++)
    if (s->cold_ok[k] && memcmp(&h, &s->cold_host[k], sizeof(TorreCold)) == 0) {
      s->cold_last = k;
      return s->cold_dev[k];
    }
  const int k = s->cold_last ^ 1;
  if (!s->cold_dev[k] && dalloc(s, &s->cold_dev[k], sizeof(TorreCold))) return nullptr;
  if (hipStreamSynchronize(s->stream) != hipSuccess ||
      hipMemcpy(s->cold_dev[k], &h, sizeof(TorreCold), hipMemcpyHostToDevice) != hipSuccess) {
    fail(DTM_ERR_DEVICE, "fused-step table upload");
    return nullptr;
  }
  memcpy(&s->cold_host[k], &h, sizeof(TorreCold));
  s->cold_ok[k] = true;
  s->cold_last = k;
  return s->cold_dev[k];
}

static int one_step(dtm_solver* s) {
  const dtm_config& c = s->cfg;
  const long long n = s->step;
  s->tf.active = (!s->adjoint && tfsf_gate(s, n)) ? 1 : 0;
  const void* F = fields_ptr(s);
  const bool rec = !s->adjoint && !s->replay;  // probes and flux record forward steps only
  if (rec && s->fs.nface > 0) launch_flux_snap(s->store, s->g, F, s->sc, s->fs, s->stream);
  const double th = (double)n * c.dt, te = ((double)n + 0.5) * c.dt;
  if (s->torre) {
    // A plane wave along x is folded into the fused step (both half-steps'
    // factor tables); other sources apply k_tfsf's H correction to
    // H^{n-1/2} before the step (exact: the step is linear) and its E
    // correction after it.
    const bool fold = s->tf.active && s->fold_base && !s->no_fold;
    Tfsf tfh{};
    if (s->tf.active) tfh = tfsf_for(s, n, 0);
    if (s->tf.active && !fold) {
      ev_begin(s, 4);
      launch_tfsf(s->store, 2, s->g, F, s->sc, s->pm, s->mt, s->ade, tfh, s->bands, th, s->stream);
      ev_end(s);
    }
    const TorreCold* cold = torre_cold(s);
    if (!cold) return DTM_ERR_DEVICE;
    ev_begin(s, 0);
    launch_torre(s->store, s->g, s->sc, s->mt, cold, &s->whole, 1, s->stream, fold ? fold_for(s, n, 0) : nullptr);
    ev_end(s);
    ev_begin(s, 4);
    launch_ade(s->store, s->g, F, fields_next_ptr(s), s->sc, s->mt, s->ade, s->stream);
    swap_buffers(s);
    F = fields_ptr(s);
    if (s->tf.active && !fold)
      launch_tfsf(s->store, 1, s->g, F, s->sc, s->pm, s->mt, s->ade, tfsf_for(s, n, 1), s->bands, te, s->stream);
  } else {
    // A plane wave along x is folded into the sweeps (k_tfsf's terms added to
    // the derivatives, from per-plane factors); other plane waves, the dipole
    // and a plane wave with the FP32 zone keep k_tfsf.
    const bool fold = s->tf.active && s->fold_base && !s->sc.xz.on && !s->no_fold;
    // H half-step: incident E at t = n dt (genAlg.py:220 with time 0)
    Tfsf tfh{};
    if (s->tf.active) tfh = tfsf_for(s, n, 0);
    ev_begin(s, 0);
    const SideStream* side = s->sc.xz.on && s->side.st ? &s->side : nullptr;
    launch_update_h(s->store, s->g, F, s->sc, s->pm, &s->whole, 1, s->stream, fold ? fold_for(s, n, 0) : nullptr,
                    side);
    ev_end(s);
    if (s->tf.active && !fold) {
      ev_begin(s, 4);
      launch_tfsf(s->store, 0, s->g, F, s->sc, s->pm, s->mt, s->ade, tfh, s->bands, th, s->stream);
      ev_end(s);
    }
    // E half-step: incident H at t = (n+1/2) dt (genAlg.py:499)
    Tfsf tfe{};
    if (s->tf.active) tfe = tfsf_for(s, n, 1);
    ev_begin(s, 2);
    launch_update_e(s->store, s->g, F, s->sc, s->pm, s->mt, s->ade, &s->whole, 1, s->stream,
                    fold ? fold_for(s, n, 1) : nullptr, side);
    ev_end(s);
    ev_begin(s, 4);
    launch_ade(s->store, s->g, F, F, s->sc, s->mt, s->ade, s->stream);
    if (s->tf.active && !fold)
      launch_tfsf(s->store, 1, s->g, F, s->sc, s->pm, s->mt, s->ade, tfe, s->bands, te, s->stream);
  }
  if (c.nsoft > 0 && !s->adjoint) {
    Soft so{};
    so.n = c.nsoft;
    const double th = ((double)n + 0.5) * c.dt;
    for (int q = 0; q < c.nsoft; q++) {
      const dtm_soft_source& ss = c.soft[q];
      const size_t idx = ((size_t)ss.i * c.ny + ss.j) * c.nz + ss.k;
      const ClassEntry& ce = s->classes[cls_of(s, idx)];
      double ie = ce.inv_eps[ss.comp];
      if (s->mt.d_on) {
        const int di = ss.i - s->mt.d_i0, dj = ss.j - s->mt.d_j0, dk = ss.k - s->mt.d_k0;
        if (di >= 0 && di < s->mt.d_ni && dj >= 0 && dj < s->mt.d_nj && dk >= 0 && dk < s->mt.d_nk)
          ie = s->design_host[((size_t)di * s->mt.d_nj + dj) * s->mt.d_nk + dk];
      }
      const double J = ss.amp * wave_value(ss.waveform, ss.w, ss.kEnv, ss.t0, ss.tw, th);
      so.comp[q] = ss.comp;
      so.q[q] = (long long)idx;
      so.ijk[q][0] = ss.i;
      so.ijk[q][1] = ss.j;
      so.ijk[q][2] = ss.k;
      so.val[q] = (float)(c.dt * ie * J);
    }
    launch_soft(s->store, s->g, F, s->sc, so, s->stream);
  }
  const int np = rec ? (int)s->probes.size() : 0;
  const int nf = rec ? s->fs.nface : 0;
  if (nf > 0) launch_flux(s->store, s->g, F, s->sc, s->fs, s->series_row, c.dt, s->stream);
  if (np > 0) {
    Probes pr{};
    pr.n = np;
    pr.desc = s->probes_dev;
    pr.series = s->series_dev;
    pr.row = s->series_row;
    pr.nfreq = s->nfreq;
    pr.omega = s->omega_dev;
    pr.dft = s->dft_dev;
    pr.dt = c.dt;
    pr.step = n;
    launch_record(s->store, s->g, F, s->sc, pr, s->stream);
  }
  ev_end(s);
  if (int rc = launch_result()) return rc;
  if (np > 0 || nf > 0) {
    s->series_row++;
    if (s->series_row == s->series_cap)
      if (int rc = flush_series(s)) return rc;
  }
  s->step++;
  return 0;
}

int dtm_step(dtm_solver* s, int nsteps) {
  if (!s || nsteps < 0) return fail(DTM_ERR_ARG, "bad step arguments");
  if (int rc = start(s)) return rc;
  for (int t = 0; t < nsteps; t++)
    if (int rc = one_step(s)) return rc;
  return launch_result();
}

long long dtm_time_step(const dtm_solver* s) { return s ? s->step : -1; }

int dtm_sync(dtm_solver* s) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int dtm_probe_count(const dtm_solver* s) { return s ? (int)s->probes.size() : 0; }

int dtm_probe_read(dtm_solver* s, float* out, int nsteps_max) {
  if (!s || (!out && nsteps_max > 0)) return fail(DTM_ERR_ARG, "null argument");
  if (int rc = flush_series(s)) return rc;
  const int np = (int)s->probes.size();
  if (np == 0) return 0;
  const long long rows = (long long)(s->series_host.size() / np);
  const long long r = std::min<long long>(rows, nsteps_max);
  memcpy(out, s->series_host.data(), sizeof(float) * (size_t)r * np);
  return (int)r;
}

int dtm_probe_last(dtm_solver* s, float* out) {
  if (!s || !out) return fail(DTM_ERR_ARG, "null argument");
  const int np = (int)s->probes.size();
  if (np == 0 || s->step == 0) return 0;
  if (s->series_row > 0) {  // the newest row is still in the device buffer
    HIPCHK(hipMemcpyAsync(out, s->series_dev + (size_t)(s->series_row - 1) * np, sizeof(float) * np,
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    if (s->series_host.size() < (size_t)np) return 0;
    memcpy(out, s->series_host.data() + s->series_host.size() - np, sizeof(float) * np);
  }
  return 1;
}

int dtm_dft_read(dtm_solver* s, double* out) {
  if (!s || !out) return fail(DTM_ERR_ARG, "null argument");
  const size_t cnt = 2 * s->probes.size() * (size_t)s->nfreq;
  if (cnt == 0) return 0;
  if (!s->dft_dev) {
    memset(out, 0, sizeof(double) * cnt);
    return 0;
  }
  HIPCHK(hipMemcpyAsync(out, s->dft_dev, sizeof(double) * cnt, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

static void* comp_ptr(dtm_solver* s, int comp) { return comp < 3 ? s->E[comp] : s->H[comp - 3]; }

// Time-stepped state of the solver: every registered buffer's contents plus
// which of the double buffers is current.
struct StateImage {
  void *E[3], *H[3], *E2[3], *H2[3];
  Cpml pm, pmn;
};
static StateImage state_ptrs(const dtm_solver* s) {
  StateImage im;
  for (int c = 0; c < 3; c++) {
    im.E[c] = s->E[c];
    im.H[c] = s->H[c];
    im.E2[c] = s->E2[c];
    im.H2[c] = s->H2[c];
  }
  im.pm = s->pm;
  im.pmn = s->pmn;
  return im;
}
static size_t state_bytes(const dtm_solver* s) {
  size_t t = 0;
  for (auto& b : s->state) t += b.second;
  return t;
}
static int save_state(dtm_solver* s, char* dst) {
  for (auto& b : s->state) {
    HIPCHK(hipMemcpyAsync(dst, b.first, b.second, hipMemcpyDeviceToDevice, s->stream));
    dst += b.second;
  }
  return 0;
}
static int load_state(dtm_solver* s, const char* src, const StateImage& im) {
  for (auto& b : s->state) {
    HIPCHK(hipMemcpyAsync(b.first, src, b.second, hipMemcpyDeviceToDevice, s->stream));
    src += b.second;
  }
  for (int c = 0; c < 3; c++) {
    s->E[c] = im.E[c];
    s->H[c] = im.H[c];
    s->E2[c] = im.E2[c];
    s->H2[c] = im.H2[c];
  }
  s->pm = im.pm;
  s->pmn = im.pmn;
  field_structs(s);
  return 0;
}

int dtm_adjoint_set_checkpoints(dtm_solver* s, int every) {
  if (!s || every < 0) return fail(DTM_ERR_ARG, "checkpoint interval must be >= 0");
  s->adj_every = every;
  return 0;
}

int dtm_adjoint_design_gradient(dtm_solver* s, int nsteps, double omega, int nmon, const int* mon_ijk,
                                double* J_out, double* grad_inv_eps) {
  if (!s || nsteps < 1 || nmon < 1 || !mon_ijk || !J_out || !grad_inv_eps)
    return fail(DTM_ERR_ARG, "bad adjoint arguments");
  const dtm_config& c = s->cfg;
  if (s->store != DTM_STORE_FP32) return fail(DTM_ERR_ARG, "the adjoint needs FP32 storage");
  if (s->ade.on) return fail(DTM_ERR_ARG, "the adjoint does not cover dispersive or gyrotropic cells");
  if (!s->mt.d_on) return fail(DTM_ERR_STATE, "no design box (dtm_set_design_eps)");
  if (s->step != 0) return fail(DTM_ERR_STATE, "the adjoint starts at step 0");
  if (s->pm.split) return fail(DTM_ERR_ARG, "the adjoint runs with the CPML, not the split-field layer");
  // The backward pass steps the adjoint fields with the forward CPML update,
  // which is the exact transpose outside the absorbing layer only: design
  // cells and monitors must lie outside every slab (adjacent is exact).
  auto in_layer = [&](int i, int j, int k) {
    const int v[3] = {i, j, k}, n[3] = {c.nx, c.ny, c.nz};
    for (int a = 0; a < 3; a++)
      if (v[a] < c.npml[2 * a] || v[a] >= n[a] - c.npml[2 * a + 1]) return true;
    return false;
  };
  if (in_layer(s->mt.d_i0, s->mt.d_j0, s->mt.d_k0) ||
      in_layer(s->mt.d_i0 + s->mt.d_ni - 1, s->mt.d_j0 + s->mt.d_nj - 1, s->mt.d_k0 + s->mt.d_nk - 1))
    return fail(DTM_ERR_ARG, "the design box reaches into the CPML (the adjoint is exact outside the layer only)");
  if (int rc = start(s)) return rc;
  const Materials& mt = s->mt;
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  std::vector<long long> q(nmon);
  std::vector<float> ie(nmon);
  bool mon_in_design = false;  // a monitor inside the design box keeps injection and gradient apart
  for (int m = 0; m < nmon; m++) {
    const int i = mon_ijk[3 * m], j = mon_ijk[3 * m + 1], k = mon_ijk[3 * m + 2];
    if (i < 0 || i >= c.nx || j < 0 || j >= c.ny || k < 0 || k >= c.nz) return fail(DTM_ERR_ARG, "monitor outside grid");
    if (in_layer(i, j, k))
      return fail(DTM_ERR_ARG, "a monitor lies in the CPML (the adjoint is exact outside the layer only)");
    q[m] = ((long long)i * c.ny + j) * c.nz + k;
    ie[m] = s->classes[cls_of(s, (size_t)q[m])].inv_eps[2];
    const int di = i - mt.d_i0, dj = j - mt.d_j0, dk = k - mt.d_k0;
    if (di >= 0 && di < mt.d_ni && dj >= 0 && dj < mt.d_nj && dk >= 0 && dk < mt.d_nk) {
      ie[m] = s->design_host[((size_t)di * mt.d_nj + dj) * mt.d_nk + dk];
      mon_in_design = true;
    }
  }
  // History mode keeps the design-box E of every step (nsteps+1 snapshots).
  // Checkpoint mode (interval K) keeps the whole state every K steps and,
  // segment by segment on the way back, recomputes the segment forward to
  // rebuild its K+1 snapshots; the adjoint state waits in a stash meanwhile.
  // The last L segments keep their snapshots from the forward pass instead
  // of a checkpoint: L = 1 + (checkpoint bytes / segment snapshot bytes), so
  // the tail never takes more memory than the checkpoints it replaces plus
  // one segment's snapshots, and the way back starts without a recompute.
  const bool ckpt = s->adj_every > 0 && s->adj_every < nsteps;
  const int K = ckpt ? s->adj_every : nsteps;
  const int nseg = ckpt ? (nsteps + K - 1) / K : 1;
  const size_t sbytes = ckpt ? state_bytes(s) : 0;
  const size_t seg_hist = sizeof(float) * 3 * (size_t)nd * (size_t)K;
  const int ntail = ckpt ? (int)std::min<size_t>((size_t)nseg, 1 + sbytes / std::max<size_t>(seg_hist, 1)) : 1;
  const int nck = nseg - ntail;  // segments rebuilt from a checkpoint
  const int n_last = nck * K;    // first step whose snapshots the forward pass keeps
  const size_t hist_bytes = sizeof(float) * 3 * (size_t)nd * ((size_t)std::max(K, nsteps - n_last) + 1);
  void *d_q = nullptr, *d_ie = nullptr, *d_ser = nullptr, *d_hist = nullptr, *d_acc = nullptr, *d_F = nullptr;
  void *d_ck = nullptr, *d_stash = nullptr;
  auto release = [&]() {
    for (void* p : {d_q, d_ie, d_ser, d_hist, d_acc, d_F, d_ck, d_stash})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc(&d_q, sizeof(long long) * nmon) != hipSuccess || hipMalloc(&d_ie, sizeof(float) * nmon) != hipSuccess ||
      hipMalloc(&d_ser, sizeof(float) * (size_t)nmon * nsteps) != hipSuccess ||
      hipMalloc(&d_hist, hist_bytes) != hipSuccess || hipMalloc(&d_acc, sizeof(double) * nd) != hipSuccess ||
      hipMalloc(&d_F, sizeof(double) * 2 * nmon) != hipSuccess ||
      (nck > 0 && (hipMalloc(&d_ck, sbytes * nck) != hipSuccess || hipMalloc(&d_stash, sbytes) != hipSuccess))) {
    (void)hipGetLastError();
    release();
    return fail(DTM_ERR_NOMEM, "adjoint temporaries (design-box history " + std::to_string(hist_bytes >> 20) +
                                   " MiB, checkpoints " + std::to_string((sbytes * (nck + 1)) >> 20) + " MiB)");
  }
  auto* hq = static_cast<long long*>(d_q);
  auto* hie = static_cast<float*>(d_ie);
  auto* ser = static_cast<float*>(d_ser);
  auto* hist = static_cast<float*>(d_hist);
  auto* acc = static_cast<double*>(d_acc);
  auto* Fd = static_cast<double*>(d_F);
  auto* ck = static_cast<char*>(d_ck);
  auto* stash = static_cast<char*>(d_stash);
  std::vector<StateImage> ck_ptrs(nck);
  const size_t snap = 3 * (size_t)nd;
  int rc = 0;
  do {
    if (hipMemcpyAsync(hq, q.data(), sizeof(long long) * nmon, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(hie, ie.data(), sizeof(float) * nmon, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemsetAsync(acc, 0, sizeof(double) * nd, s->stream) != hipSuccess) {
      rc = fail(DTM_ERR_DEVICE, "adjoint set-up copies");
      break;
    }
    // forward: record Ez at the monitors every step, and the design-box E
    // every step (history) or the whole state every K steps (checkpoints)
    // plus the tail segments' snapshots
    for (int n = 0; n < nsteps && !rc; n++) {
      if (n < n_last && n % K == 0) {
        ck_ptrs[n / K] = state_ptrs(s);
        if ((rc = save_state(s, ck + (size_t)(n / K) * sbytes))) break;
      }
      if (n == n_last) launch_adj_snap(s->g, fields_ptr(s), mt, hist, s->stream);
      rc = one_step(s);
      launch_adj_monitor(fields_ptr(s), hq, nmon, ser + (size_t)n * nmon, s->stream);
      if (n >= n_last) launch_adj_snap(s->g, fields_ptr(s), mt, hist + (size_t)(n - n_last + 1) * snap, s->stream);
    }
    if (!rc) rc = launch_result();
    if (rc) break;
    std::vector<float> hs((size_t)nmon * nsteps);
    if (hipMemcpyAsync(hs.data(), ser, sizeof(float) * hs.size(), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      rc = fail(DTM_ERR_DEVICE, "adjoint series copy");
      break;
    }
    // F_m = sum_n Ez_m^{n+1} e^{i w (n+1) dt} dt, J = sum_m |F_m|^2
    std::vector<double> F(2 * (size_t)nmon, 0.0);
    for (int n = 0; n < nsteps; n++) {
      const double t = (double)(n + 1) * c.dt;
      const double cs = cos(omega * t) * c.dt, sn = sin(omega * t) * c.dt;
      for (int m = 0; m < nmon; m++) {
        F[2 * m] += hs[(size_t)n * nmon + m] * cs;
        F[2 * m + 1] += hs[(size_t)n * nmon + m] * sn;
      }
    }
    double J = 0.0;
    for (int m = 0; m < nmon; m++) J += F[2 * m] * F[2 * m] + F[2 * m + 1] * F[2 * m + 1];
    *J_out = J;
    if (hipMemcpyAsync(Fd, F.data(), sizeof(double) * 2 * nmon, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
      rc = fail(DTM_ERR_DEVICE, "adjoint F copy");
      break;
    }
    // adjoint: zeta = M lambda_E, eta = -lambda_H stepped backward by the
    // forward sweeps; zeta^N = M dJ/dE^N
    if ((rc = zero_state(s))) break;
    s->adjoint = true;
    launch_adj_inject(fields_ptr(s), hq, hie, Fd, nmon, omega, c.dt, nsteps, s->stream);
    for (int seg = nseg - 1; seg >= 0 && !rc; seg--) {
      const int n0 = seg * K, n1 = std::min(n0 + K, nsteps);
      // this segment's snapshots: kept from the forward pass (tail) or
      // rebuilt at the start of the buffer (the tail is done with by then)
      float* const hs0 = hist + (seg >= nck ? (size_t)(n0 - n_last) * snap : 0);
      if (seg < nck) {  // rebuild the segment's history from its checkpoint, the adjoint state stashed
        const StateImage adj = state_ptrs(s);
        const long long adj_step = s->step;
        if ((rc = save_state(s, stash)) || (rc = load_state(s, ck + (size_t)seg * sbytes, ck_ptrs[seg]))) break;
        s->adjoint = false;
        s->replay = true;
        s->step = n0;
        s->tab_n0 = -1;
        launch_adj_snap(s->g, fields_ptr(s), mt, hist, s->stream);
        for (int n = n0; n < n1 && !rc; n++) {
          rc = one_step(s);
          launch_adj_snap(s->g, fields_ptr(s), mt, hist + (size_t)(n - n0 + 1) * snap, s->stream);
        }
        s->replay = false;
        s->adjoint = true;
        if (rc || (rc = load_state(s, stash, adj))) break;
        s->step = adj_step;
      }
      // gradient of step n, one backward step, injection at step n; the
      // injection and the next step's gradient share a launch when they
      // touch disjoint cells
      launch_adj_grad(s->g, fields_ptr(s), mt, hs0 + (size_t)(n1 - 1 - n0) * snap, hs0 + (size_t)(n1 - n0) * snap,
                      acc, s->stream);
      for (int n = n1 - 1; n >= n0 && !rc; n--) {
        rc = one_step(s);
        const bool next = n - 1 >= n0;  // the gradient of step n-1 follows in this segment
        if (n >= 1 && next && !mon_in_design) {
          launch_adj_inject_grad(s->g, fields_ptr(s), mt, hs0 + (size_t)(n - 1 - n0) * snap,
                                 hs0 + (size_t)(n - n0) * snap, acc, hq, hie, Fd, nmon, omega, c.dt, n, s->stream);
          continue;
        }
        if (n >= 1) launch_adj_inject(fields_ptr(s), hq, hie, Fd, nmon, omega, c.dt, n, s->stream);
        if (next)
          launch_adj_grad(s->g, fields_ptr(s), mt, hs0 + (size_t)(n - 1 - n0) * snap, hs0 + (size_t)(n - n0) * snap,
                          acc, s->stream);
      }
    }
    s->adjoint = false;
    s->replay = false;
    if (!rc) rc = launch_result();
    if (rc) break;
    if (hipMemcpyAsync(grad_inv_eps, acc, sizeof(double) * nd, hipMemcpyDeviceToHost, s->stream) != hipSuccess) {
      rc = fail(DTM_ERR_DEVICE, "adjoint gradient copy");
      break;
    }
  } while (false);
  s->adjoint = false;
  s->replay = false;
  // back to step 0 with zero fields
  if (!rc) rc = zero_state(s);
  s->step = 0;
  s->tab_n0 = -1;
  if (hipStreamSynchronize(s->stream) != hipSuccess && !rc) rc = fail(DTM_ERR_DEVICE, "adjoint sync");
  release();
  return rc;
}

// The FP32 zone's values of component comp into (to_host) or from a host
// array of the whole grid.
static int zone_copy(dtm_solver* s, int comp, float* buf, bool to_host) {
  const Exact& x = s->sc.xz;
  if (!x.on) return 0;
  const dtm_config& c = s->cfg;
  const int ni = x.i1 - x.i0, nj = x.j1 - x.j0, nk = x.k1 - x.k0;
  std::vector<float> z((size_t)ni * nj * nk);
  if (!to_host)
    for (int i = 0; i < ni; i++)
      for (int j = 0; j < nj; j++)
        memcpy(&z[((size_t)i * nj + j) * nk], &buf[((size_t)(x.i0 + i) * c.ny + (x.j0 + j)) * c.nz + x.k0],
               sizeof(float) * nk);
  const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(to_host ? (void*)z.data() : (void*)x.f[comp], to_host ? (const void*)x.f[comp] : z.data(),
                        sizeof(float) * z.size(), kind, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (to_host)
    for (int i = 0; i < ni; i++)
      for (int j = 0; j < nj; j++)
        memcpy(&buf[((size_t)(x.i0 + i) * c.ny + (x.j0 + j)) * c.nz + x.k0], &z[((size_t)i * nj + j) * nk],
               sizeof(float) * nk);
  return 0;
}

int dtm_field_read(dtm_solver* s, int comp, float* buf) {
  if (!s || !buf || comp < 0 || comp > 5) return fail(DTM_ERR_ARG, "bad field read arguments");
  const size_t n = (size_t)s->cfg.nx * s->cfg.ny * s->cfg.nz;
  if (s->store == 0) {
    HIPCHK(hipMemcpyAsync(buf, comp_ptr(s, comp), sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
  } else {
    if (!s->tmp_dev)
      if (int rc = dalloc(s, &s->tmp_dev, sizeof(float) * n)) return rc;
    const float inv = comp < 3 ? s->sc.ie[comp] : s->sc.ih[comp - 3];
    launch_to_float(1, comp_ptr(s, comp), inv, s->tmp_dev, (long long)n, s->stream);
    if (int rc = launch_result()) return rc;
    HIPCHK(hipMemcpyAsync(buf, s->tmp_dev, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return zone_copy(s, comp, buf, true);
}

// ------------------------------------------------------ save / resume
// The time-stepped state in a fixed order: the current E and H, the current
// CPML psi, the split-field parts and the ADE state (E^{n-1}, J per pole,
// plasma current).  The fused step's second buffers are not state: every
// step overwrites them completely.
static std::vector<std::pair<void*, size_t>> canonical_state(const dtm_solver* s) {
  std::unordered_map<const void*, size_t> sz;
  for (auto& b : s->state) sz[b.first] = b.second;
  std::vector<std::pair<void*, size_t>> out;
  auto add = [&](void* p) {
    auto it = sz.find(p);
    if (p && it != sz.end()) out.emplace_back(p, it->second);
  };
  for (int c = 0; c < 3; c++) add(s->E[c]);
  for (int c = 0; c < 3; c++) add(s->H[c]);
  for (int a = 0; a < 3; a++)
    for (int sl = 0; sl < 2; sl++) {
      add(s->pm.psiE[a][sl]);
      add(s->pm.psiH[a][sl]);
    }
  for (int c = 0; c < 3; c++) {
    add(s->pm.p2E[c]);
    add(s->pm.p2H[c]);
  }
  for (int c = 0; c < 3; c++) {
    add(s->ade.Em[c]);
    for (int p = 0; p < kMaxPoles; p++) {
      add(s->ade.Jp[c][p]);
      add(s->ade.Jm[c][p]);
    }
    add(s->ade.Jg[c]);
  }
  for (int c = 0; c < 6; c++) add(s->sc.xz.f[c]);  // the FP32 zone
  return out;
}

namespace {
struct StateHeader {
  char magic[8];  // "DTMSTAT1"
  int nx, ny, nz, storage, pml_kind, nbuf;
  long long step;
  long long dft_values, flux_values;
};
}  // namespace

int dtm_save_state(dtm_solver* s, const char* path) {
  if (!s || !path) return fail(DTM_ERR_ARG, "null argument");
  if (int rc = start(s)) return rc;
  const auto st = canonical_state(s);
  StateHeader h{};
  memcpy(h.magic, "DTMSTAT1", 8);
  h.nx = s->cfg.nx;
  h.ny = s->cfg.ny;
  h.nz = s->cfg.nz;
  h.storage = s->store;
  h.pml_kind = s->cfg.pml_kind;
  h.nbuf = (int)st.size();
  h.step = s->step;
  h.dft_values = s->dft_dev ? 2 * (long long)s->probes.size() * s->nfreq : 0;
  h.flux_values = s->fs.total ? s->fs.nface : 0;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(DTM_ERR_ARG, std::string("cannot open ") + path + " for writing");
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
  std::vector<char> buf;
  for (auto& b : st) {
    const unsigned long long n = b.second;
    buf.resize(b.second);
    if (hipMemcpyAsync(buf.data(), b.first, b.second, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      fclose(f);
      return fail(DTM_ERR_DEVICE, "state read-back");
    }
    ok = ok && fwrite(&n, sizeof(n), 1, f) == 1 && fwrite(buf.data(), 1, b.second, f) == b.second;
  }
  auto dump = [&](const double* dev, long long cnt) {
    if (cnt <= 0) return;
    std::vector<double> v((size_t)cnt);
    if (hipMemcpy(v.data(), dev, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    ok = ok && fwrite(v.data(), sizeof(double), (size_t)cnt, f) == (size_t)cnt;
  };
  dump(s->dft_dev, h.dft_values);
  dump(s->fs.total, h.flux_values);
  if (fclose(f) != 0) ok = false;
  return ok ? 0 : fail(DTM_ERR_DEVICE, std::string("writing ") + path + " failed");
}

int dtm_load_state(dtm_solver* s, const char* path) {
  if (!s || !path) return fail(DTM_ERR_ARG, "null argument");
  if (int rc = start(s)) return rc;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(DTM_ERR_ARG, std::string("cannot open ") + path);
  StateHeader h{};
  const auto st = canonical_state(s);
  const long long ndft = s->dft_dev ? 2 * (long long)s->probes.size() * s->nfreq : 0;
  const long long nflux = s->fs.total ? s->fs.nface : 0;
  if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, "DTMSTAT1", 8) != 0) {
    fclose(f);
    return fail(DTM_ERR_ARG, std::string(path) + " is not a saved solver state");
  }
  if (h.nx != s->cfg.nx || h.ny != s->cfg.ny || h.nz != s->cfg.nz || h.storage != s->store ||
      h.pml_kind != s->cfg.pml_kind || h.nbuf != (int)st.size() || h.dft_values != ndft || h.flux_values != nflux) {
    fclose(f);
    return fail(DTM_ERR_ARG, "saved state does not match this solver's grid, storage, layer, materials, probes "
                             "or flux boxes");
  }
  std::vector<char> buf;
  for (auto& b : st) {
    unsigned long long n = 0;
    buf.resize(b.second);
    if (fread(&n, sizeof(n), 1, f) != 1 || n != b.second || fread(buf.data(), 1, b.second, f) != b.second) {
      fclose(f);
      return fail(DTM_ERR_ARG, "saved state buffer sizes do not match this solver");
    }
    if (hipMemcpy(b.first, buf.data(), b.second, hipMemcpyHostToDevice) != hipSuccess) {
      fclose(f);
      return fail(DTM_ERR_DEVICE, "state upload");
    }
  }
  auto restore = [&](double* dev, long long cnt) -> bool {
    if (cnt <= 0) return true;
    std::vector<double> v((size_t)cnt);
    return fread(v.data(), sizeof(double), (size_t)cnt, f) == (size_t)cnt &&
           hipMemcpy(dev, v.data(), sizeof(double) * (size_t)cnt, hipMemcpyHostToDevice) == hipSuccess;
  };
  const bool ok = restore(s->dft_dev, ndft) && restore(s->fs.total, nflux);
  fclose(f);
  if (!ok) return fail(DTM_ERR_ARG, "saved state is truncated");
  if (int rc = flush_series(s)) return rc;
  s->series_host.clear();
  s->flux_host.clear();
  s->step = h.step;
  s->tab_n0 = -1;
  return 0;
}

// One component as an aivlib .arr file (the layout SeismoDrops::drop writes,
// drop.cu:84-107): int 12, three zero ints (a 12-byte comment), int dim = 3,
// int sizeof(float), the extents fastest first (nz, ny, nx), then the values
// with z fastest.
int dtm_write_arr(dtm_solver* s, int comp, const char* path) {
  if (!s || !path || comp < 0 || comp > 5) return fail(DTM_ERR_ARG, "bad .arr arguments");
  const size_t n = (size_t)s->cfg.nx * s->cfg.ny * s->cfg.nz;
  std::vector<float> v(n);
  if (int rc = dtm_field_read(s, comp, v.data())) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(DTM_ERR_ARG, std::string("cannot open ") + path + " for writing");
  const int head[9] = {12, 0, 0, 0, 3, (int)sizeof(float), s->cfg.nz, s->cfg.ny, s->cfg.nx};
  bool ok = fwrite(head, sizeof(int), 9, f) == 9 && fwrite(v.data(), sizeof(float), n, f) == n;
  if (fclose(f) != 0) ok = false;
  return ok ? 0 : fail(DTM_ERR_DEVICE, std::string("writing ") + path + " failed");
}

int dtm_field_write(dtm_solver* s, int comp, const float* buf) {
  if (!s || !buf || comp < 0 || comp > 5) return fail(DTM_ERR_ARG, "bad field write arguments");
  const size_t n = (size_t)s->cfg.nx * s->cfg.ny * s->cfg.nz;
  if (s->store == 0) {
    HIPCHK(hipMemcpyAsync(comp_ptr(s, comp), buf, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
  } else {
    if (!s->tmp_dev)
      if (int rc = dalloc(s, &s->tmp_dev, sizeof(float) * n)) return rc;
    HIPCHK(hipMemcpyAsync(s->tmp_dev, buf, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
    const float sc = comp < 3 ? s->sc.se[comp] : s->sc.sh[comp - 3];
    launch_from_float(1, s->tmp_dev, sc, comp_ptr(s, comp), (long long)n, s->stream);
    if (int rc = launch_result()) return rc;
    if (int rc = zone_copy(s, comp, const_cast<float*>(buf), false)) return rc;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

static int box_ok(const dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1) {
  const dtm_config& c = s->cfg;
  return i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 < i1 && j0 < j1 && k0 < k1 && i1 <= c.nx && j1 <= c.ny && k1 <= c.nz;
}

int dtm_energy_box(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1, double* out) {
  if (!s || !out) return fail(DTM_ERR_ARG, "null argument");
  if (!box_ok(s, i0, i1, j0, j1, k0, k1)) return fail(DTM_ERR_ARG, "energy box outside the grid or empty");
  const int ni = i1 - i0;
  if (ni > s->energy_part_cap) {  // one FP64 partial per x plane of the box
    if (int rc = dalloc(s, &s->energy_part, sizeof(double) * (size_t)ni)) return rc;
    s->energy_part_cap = ni;
  }
  const dtm_config& c = s->cfg;
  launch_energy_box(s->store, s->g, fields_ptr(s), s->sc, s->mt, Region{i0, i1, j0, j1, k0, k1}, s->energy_part,
                    c.dx * c.dy * c.dz, s->energy_dev, s->stream);
  if (int rc = launch_result()) return rc;
  HIPCHK(hipMemcpyAsync(out, s->energy_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int dtm_energy_box_noalloc(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1, double* out) {
  if (!s || !out) return fail(DTM_ERR_ARG, "null argument");
  if (!box_ok(s, i0, i1, j0, j1, k0, k1)) return fail(DTM_ERR_ARG, "energy box outside the grid or empty");
  const dtm_config& c = s->cfg;
  const int rc = energy_box_static(s->dev, s->store, s->g, fields_ptr(s), s->sc, s->mt, Region{i0, i1, j0, j1, k0, k1},
                                   c.dx * c.dy * c.dz, out, s->stream);
  if (rc == -2)
    return fail(DTM_ERR_LIMIT, "dtm_energy_box_noalloc: the box spans more than " +
                                   std::to_string(static_scratch_planes()) + " x planes (its static scratch)");
  if (int lr = launch_result()) return lr;
  if (rc) return fail(DTM_ERR_DEVICE, "dtm_energy_box_noalloc: device error");
  return 0;
}

int dtm_energy(dtm_solver* s, double* out) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  return dtm_energy_box(s, 0, s->cfg.nx, 0, s->cfg.ny, 0, s->cfg.nz, out);
}

int dtm_add_flux_box(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  if (s->started) return fail(DTM_ERR_STATE, "flux boxes must be added before the first step");
  if ((int)s->flux_faces.size() >= 6 * kMaxFluxBoxes) return fail(DTM_ERR_LIMIT, "too many flux boxes");
  const dtm_config& c = s->cfg;
  const int lo[3] = {i0, j0, k0}, hi[3] = {i1, j1, k1}, n[3] = {c.nx, c.ny, c.nz};
  const double d[3] = {c.dx, c.dy, c.dz};
  // every face plane needs the E cells on both sides and the H cell on it
  for (int a = 0; a < 3; a++)
    if (lo[a] < 1 || hi[a] <= lo[a] || hi[a] > n[a] - 1)
      return fail(DTM_ERR_ARG, "flux box faces must lie on planes 1 .. n-1 of every axis, low < high");
  const int id = (int)s->flux_faces.size() / 6;
  for (int a = 0; a < 3; a++) {
    const int b = (a + 1) % 3, e = (a + 2) % 3;
    for (int side = 0; side < 2; side++) {
      FluxFace f{};
      f.a = a;
      f.pa = side ? hi[a] : lo[a];
      f.b0 = lo[b];
      f.b1 = hi[b];
      f.c0 = lo[e];
      f.c1 = hi[e];
      const long long nb = hi[b] - lo[b], nc = hi[e] - lo[e];
      const long long nodes = (nb + 1) * nc + nb * (nc + 1);
      if (nodes + s->flux_snap_n > 0x7fffffffLL) return fail(DTM_ERR_LIMIT, "flux faces too large");
      f.nodes = (int)nodes;
      f.snap0 = s->flux_snap_n;
      s->flux_snap_n += (int)nodes;
      f.area = d[b] * d[e];
      f.sign = side ? 1.0 : -1.0;
      s->flux_faces.push_back(f);
    }
  }
  return id;
}

int dtm_set_tfsf_fold(dtm_solver* s, int on) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  s->no_fold = !on;
  return 0;
}

int dtm_set_fp32_zone(dtm_solver* s, int i0, int i1, int j0, int j1, int k0, int k1) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  if (s->started) return fail(DTM_ERR_STATE, "the FP32 zone must be set before the first step");
  if (s->store != DTM_STORE_FP16) return fail(DTM_ERR_ARG, "the FP32 zone applies to FP16 storage only");
  if (s->torre) return fail(DTM_ERR_ARG, "the FP32 zone runs with the two-pass sweeps only");
  if (s->sc.xz.on) return fail(DTM_ERR_STATE, "the FP32 zone is already set");
  if (wide_grid(s->g)) return fail(DTM_ERR_LIMIT, "the FP32 zone is limited to grids below 2^32 cells");
  if (!box_ok(s, i0, i1, j0, j1, k0, k1)) return fail(DTM_ERR_ARG, "FP32 zone outside the grid or empty");
  const long long cells = (long long)(i1 - i0) * (j1 - j0) * (k1 - k0);
  if (cells > kMaxZoneCells) return fail(DTM_ERR_LIMIT, "FP32 zone larger than 2^27 cells");
  Exact x{};
  for (int c = 0; c < 6; c++)
    if (int rc = sdalloc(s, &x.f[c], sizeof(float) * (size_t)cells)) return rc;
  if (!s->side.st) {  // a second stream (default priority) for the zone's column
    HIPCHK(hipStreamCreate(&s->side.st));
    HIPCHK(hipEventCreateWithFlags(&s->side.fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&s->side.join, hipEventDisableTiming));
  }
  x.i0 = i0;
  x.i1 = i1;
  x.j0 = j0;
  x.j1 = j1;
  x.k0 = k0;
  x.k1 = k1;
  x.on = 1;
  s->sc.xz = x;
  return 0;
}

int dtm_flux_read(dtm_solver* s, double* out, int nsteps_max) {
  if (!s || (!out && nsteps_max > 0)) return fail(DTM_ERR_ARG, "null argument");
  if (int rc = flush_series(s)) return rc;
  const int nf = (int)s->flux_faces.size();
  if (nf == 0) return 0;
  const long long rows = (long long)(s->flux_host.size() / nf);
  const long long r = std::min<long long>(rows, nsteps_max);
  memcpy(out, s->flux_host.data(), sizeof(double) * (size_t)r * nf);
  return (int)r;
}

int dtm_flux_total(dtm_solver* s, double* out) {
  if (!s || !out) return fail(DTM_ERR_ARG, "null argument");
  const int nf = (int)s->flux_faces.size();
  if (nf == 0) return 0;
  if (!s->fs.total) {
    memset(out, 0, sizeof(double) * nf);
    return 0;
  }
  HIPCHK(hipMemcpyAsync(out, s->fs.total, sizeof(double) * nf, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int dtm_time_steps(dtm_solver* s, int nsteps, double* ms) {
  if (!s || !ms || nsteps < 0) return fail(DTM_ERR_ARG, "bad timing arguments");
  if (int rc = start(s)) return rc;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, s->stream));
  for (int t = 0; t < nsteps; t++)
    if (int rc = one_step(s)) return rc;
  HIPCHK(hipEventRecord(b, s->stream));
  HIPCHK(hipEventSynchronize(b));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = f;
  return 0;
}

const char* dtm_kernel_name(const dtm_solver* s) {
  if (s && s->torre) return s->store == 0 ? "k_torre2<fp32>" : "k_torre2<fp16>";
  return kernel_variant_name(s ? s->store : 0);
}

int dtm_set_kernel_timing(dtm_solver* s, int on) {
  if (!s) return fail(DTM_ERR_ARG, "null solver");
  s->timing = on != 0;
  return 0;
}

int dtm_kernel_times(dtm_solver* s, double* ms, long long* count) {
  if (!s || !ms || !count) return fail(DTM_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(s->stream));
  for (size_t q = 0; q < s->ev_used; q++) {
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, s->ev_pool[q][0], s->ev_pool[q][1]));
    s->grp_ms[s->ev_group[q]] += f;
    s->grp_cnt[s->ev_group[q]] += 1;
  }
  s->ev_used = 0;
  for (int g = 0; g < 5; g++) {
    ms[g] = s->grp_ms[g];
    count[g] = s->grp_cnt[g];
    s->grp_ms[g] = 0;
    s->grp_cnt[g] = 0;
  }
  return 0;
}

long long dtm_group_cells(const dtm_solver* s, int g) {
  if (!s) return -1;
  auto cells = [](const Region& r) {
    return (long long)std::max(0, r.i1 - r.i0) * std::max(0, r.j1 - r.j0) * std::max(0, r.k1 - r.k0);
  };
  if (g < 0 || g > 3) return 0;
  if (s->torre) return g == 0 ? cells(s->whole) : 0;
  return (g == 0 || g == 2) ? cells(s->whole) : 0;
}

}  // extern "C"
