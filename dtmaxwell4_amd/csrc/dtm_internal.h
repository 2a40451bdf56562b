// dtm_internal.h — POD parameter blocks passed by value to the gfx950 kernels,
// and the launcher entry points implemented in dtm_kernels.hip.
//
// Device layout (HBM): six field arrays, SoA, C order [nx][ny][nz] with z
// fastest so a 64-lane wavefront reads 64 consecutive z cells of one (x,y)
// row.  This replaces the reference's DiamondRag rag structs
// (params.h:242-262) and their copyrags.cuh shuffles.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtm {

constexpr int kMaxClasses = 256;
constexpr int kMaxPoles = 4;
constexpr int kMaxProbes = 4096;

struct Grid {
  int nx, ny, nz;
  long long sx, sy;  // strides of x and y (sz = 1)
  int per[3];        // periodic axes
  float rd[3];       // 1/(24 d_a): the (2,4) weights 27/24, -1/24 (defs.h:3-13)
  float dt;
  const void* zero;  // 256 zero bytes
  unsigned zidx;     // = nx*ny*nz: index of the zero pad that ends every field and class array
  long long zidxw;   // the same index in 64 bits (grids of 2^32 - kPad cells and more)
};

// Every field array and the class array hold kPad zero elements after the
// last cell; loads of positions outside the grid read index nx*ny*nz.
constexpr int kPad = 64;
// Grids whose cell indices need 64 bits (the sweeps' wide variants).
inline bool wide_grid(const Grid& g) { return g.zidxw > 0xffffffffLL - kPad; }

// CPML: psi lives only in the slabs.  For axis a the two affected
// components are (a+1)%3 (slot 0) and (a+2)%3 (slot 1).
struct Cpml {
  int on;
  int plo[3], phi[3];
  const float* bE[3];  // indexed by cell index along the axis (E: odd half-coords)
  const float* cE[3];
  const float* bH[3];  // (H: even half-coords)
  const float* cH[3];
  float* psiE[3][2];
  float* psiH[3][2];
  // split-field layer: b* hold k1 = 2 k2 - 1 and c* hold k2 = 1/(1 + sigma dt/2)
  // (1 and 1 outside the slab); p2* hold, per component and cell, the part
  // along the component's second curl axis (the first is field - p2)
  int split;
  float* p2E[3];
  float* p2H[3];
};

// Material class table: class byte per cell -> per-component inverse eps and
// "special" (dispersive / gyrotropic) flags.  Replaces get_mat_indev and the
// texture lookups of defs.h:128-182.
struct ClassEntry {
  float inv_eps[3];
  uint8_t mat[3];
  uint8_t special;  // bit c set: component c uses the ADE path; bit 3: gyro
};

struct Materials {
  const uint8_t* cls;  // nullptr => all cells are class 0
  const ClassEntry* table;
  int nclass;
  int cx0, cx1;        // the x planes that hold cells of a class other than 0: [cx0, cx1)
  // continuous-eps design box (inverse design)
  int d_on, d_i0, d_j0, d_k0, d_ni, d_nj, d_nk;
  const float* d_inv_eps;
};

// Lorentz/Drude ADE (dispersion.cpp:41-99) over a bounding box of cells.
struct AdeMat {
  int np;
  float kE[3];
  float kEJ[kMaxPoles][2];
  float kJ[kMaxPoles][5];
};
struct GyroMat {
  float A[9], B[9];
};
struct Ade {
  int on;
  int i0, j0, k0, ni, nj, nk;  // bounding box
  long long bsx, bsy;
  float* Em[3];
  float* Jp[3][kMaxPoles];
  float* Jm[3][kMaxPoles];
  const AdeMat* mats;  // indexed by material id
  int gyro_on;
  float* Jg[3];
  const GyroMat* gmats;
  float* curl[3];  // curl H of special cells, written by the E sweep
};

// TF/SF: integer inSF bounds on half-coordinates per axis (computed on the
// host with the same double predicate as signal.hpp:153-160) and the
// incident-field parameters (double).
struct Tfsf {
  int active;
  int kind, waveform, sf_inside;
  int hlo[3], hhi[3];  // inside <=> hlo <= h <= hhi on every axis
  // cells that can see a face: [clo,chi] on every axis, face bands
  // [fl0,fl1] and [fh0,fh1] along the face normal
  int clo[3], chi[3], fl0[3], fl1[3], fh0[3], fh1[3];
  double half_d[3];    // d_a / 2
  double w, kEnv, cSrc, Imp, t0, tw, amp;
  double st, ct, sph, cph, sps, cps;  // plane wave angles
  double kx, ky, kz, pos[3];
  double dip[3];  // dipole position
  // plane wave with k along one axis: the incident component c at
  // half-coordinate h is P[c] * table[h_a - tab_h0], with table the
  // waveform amp*f0(phase) of the current half-step along axis tab_axis
  int tab_axis;  // -1: evaluate per point
  int tab_h0, tab_n;
  const double* table;
  double P[6];
  // a dipole with a small box: the incident value of component c at
  // half-coordinates h in [ptab_h0, ptab_h0 + ptab_n) on every axis, filled
  // before each correction launch (k_tfsf_points), index ((c * n0 + h0) *
  // n1 + h1) * n2 + h2 relative to ptab_h0; null: evaluate per point
  double* ptab;
  int ptab_h0[3], ptab_n[3];
};
// Point tables up to 2^20 values (boxes up to about 24 cells a side): their
// volume stays near the face bands' own evaluation count.
constexpr long long kMaxPointTable = 1LL << 20;

template <int STORE> struct Fields;
template <> struct Fields<0> {
  float* E[3];
  float* H[3];
};
template <> struct Fields<1> {
  __half* E[3];
  __half* H[3];
};

// FP16 storage with an FP32 zone: a box of cells whose six components are
// also kept in FP32.  Inside it every kernel reads and writes the FP32 copy;
// the FP16 arrays receive its rounding, which is what cells outside the box
// read.  Used around a dipole source, where the total field just outside
// its TF/SF box is large and its FP16 rounding feeds the scattered field.
struct Exact {
  int on;
  int i0, i1, j0, j1, k0, k1;
  float* f[6];  // Ex .. Hz, index ((i - i0) * (j1 - j0) + (j - j0)) * (k1 - k0) + (k - k0)
};
constexpr long long kMaxZoneCells = 1LL << 27;
__host__ __device__ __forceinline__ int exact_idx(const Exact& x, int i, int j, int k) {
  if (!x.on || i < x.i0 || i >= x.i1 || j < x.j0 || j >= x.j1 || k < x.k0 || k >= x.k1) return -1;
  return ((i - x.i0) * (x.j1 - x.j0) + (j - x.j0)) * (x.k1 - x.k0) + (k - x.k0);
}

// ------------------------------------------- TF/SF folded into the sweeps
// A plane wave along x (table axis 0): the incident value at every stencil
// neighbour along y or z of a component is P[cq] T(h_x) with h_x the
// component's own x half-coordinate, so tfsf_dcorr along y or z separates into
// a per-plane factor P[cq] T(h_x) [h_x inside the box] and a per-lane integer
// weight W(h) = sum over the neighbours across the face of wt * (+-1).  Along
// x it is a per-plane sum X(h_x) = sum wt * (+-1) T(h_x + o) times a per-lane
// inside flag.  k_tfsf_fold writes the six per-plane factors of each half-step
// (rd_a folded in); the two-pass sweeps and the fused step add
// weight * factor  to the six derivatives.
__host__ __device__ __forceinline__ int fold_w(int h, int lo, int hi, int sf_inside) {
  const bool pin = h >= lo && h <= hi;
  int w = 0;
  const int ho[4] = {-3, -1, 1, 3}, wt[4] = {1, -27, 27, -1};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool qin = h + ho[q] >= lo && h + ho[q] <= hi;
    if (qin == pin) continue;
    w += ((sf_inside ? qin : !qin) ? wt[q] : -wt[q]);
  }
  return w;  // |w| <= 28
}
// Six weights of one cell and half-step, packed: four signed 6-bit fields
// (bits 0, 6, 12, 18: factors 0, 1, 2, 5) and two flags (bits 24, 25:
// factors 3, 4).  H half-step (factors: Hx/y, Hx/z, Hy/z, Hy/x, Hz/x, Hz/y):
//   W_y(2j) zin(2k), W_z(2k) yin(2j), W_z(2k) yin(2j+1), yin(2j+1) zin(2k),
//   yin(2j) zin(2k+1), W_y(2j) zin(2k+1);
// E half-step (Ex/y, Ex/z, Ey/z, Ey/x, Ez/x, Ez/y):
//   W_y(2j+1) zin(2k+1), W_z(2k+1) yin(2j+1), W_z(2k+1) yin(2j), yin(2j) zin(2k+1),
//   yin(2j+1) zin(2k), W_y(2j+1) zin(2k).
// hd: hlo_y, hhi_y, hlo_z, hhi_z, sf_inside.
__host__ __device__ __forceinline__ int fold_pack(int half, int j, int k, const int* hd) {
  auto yin = [&](int h) { return h >= hd[0] && h <= hd[1] ? 1 : 0; };
  auto zin = [&](int h) { return h >= hd[2] && h <= hd[3] ? 1 : 0; };
  auto Wy = [&](int h) { return fold_w(h, hd[0], hd[1], hd[4]); };
  auto Wz = [&](int h) { return fold_w(h, hd[2], hd[3], hd[4]); };
  const int y0 = 2 * j, z0 = 2 * k;
  int f0, f1, f2, f3, f4, f5;
  if (half == 0) {
    f0 = Wy(y0) * zin(z0);
    f1 = Wz(z0) * yin(y0);
    f2 = Wz(z0) * yin(y0 + 1);
    f3 = yin(y0 + 1) * zin(z0);
    f4 = yin(y0) * zin(z0 + 1);
    f5 = Wy(y0) * zin(z0 + 1);
  } else {
    f0 = Wy(y0 + 1) * zin(z0 + 1);
    f1 = Wz(z0 + 1) * yin(y0 + 1);
    f2 = Wz(z0 + 1) * yin(y0);
    f3 = yin(y0) * zin(z0 + 1);
    f4 = yin(y0 + 1) * zin(z0);
    f5 = Wy(y0 + 1) * zin(z0);
  }
  return (f0 & 63) | ((f1 & 63) << 6) | ((f2 & 63) << 12) | ((f5 & 63) << 18) | (f3 << 24) | (f4 << 25);
}
__device__ __forceinline__ float fold_f(int w, int off) {  // signed 6-bit field at bit off
  return (float)((int)((unsigned)w << (26 - off)) >> 26);
}
__device__ __forceinline__ float fold_b(int w, int bit) { return (float)((w >> bit) & 1); }
// Floats per plane in the factor table; the table of one half-step starts
// with kFoldStride ints of header (hlo_y, hhi_y, hlo_z, hhi_z, sf_inside).
constexpr int kFoldStride = 8;

struct Scales {  // FP16 storage: value = stored * inv, stored = value * scale, and the FP32 zone
  float se[3], sh[3], ie[3], ih[3];
  Exact xz;
};

struct Region {  // a box of cells [i0,i1) x [j0,j1) x [k0,k1)
  int i0, i1, j0, j1, k0, k1;
};

// Regions of the fused torre kernel: per region, tiles of 61 (z) x 13 (y)
// owned cells and x segments of `seg` planes.
// Fused step: the z-CPML psi pointers and the slab bounds.
struct TorrePml {
  int on;
  int plo[3], phi[3];
  const float* hO[2];
  float* hN[2];
  const float* eO[2];
  float* eN[2];
};
// Fused step: the array pointers of one step (old / new fields Ex Ey Ez Hx
// Hy Hz, psi, materials, ADE).  The kernel reads them from device memory
// through a pointer made opaque on every plane: they are reloaded with a few
// scalar loads per plane instead of being held in scalar registers across the
// plane loop (which spills them to vector-register lanes).
struct TorreCold {
  const void* fo[6];
  void* fn[6];
  // the same arrays as one base each: component c at byte c * cbytes
  const char* fob;
  char* fnb;
  unsigned cbytes;
  TorrePml zp;
  Cpml pmo, pmn;  // old / new psi (x and y slabs)
  Materials mt;
  Ade ade;
};

struct TorreSet {
  int n;
  Region r[6];
  int nbk[6], nbj[6], nbs[6];
  int start[7];
  int total, per, seg;
};

// TF/SF face bands: for each axis a, up to two boxes of cells whose stencil
// along a can cross a face of the box (the other axes span the box).
struct TfsfBands {
  int n;
  Region r[6];
  int axis[6];
  unsigned start[7];  // first thread of each band (filled at launch)
  unsigned ovl[6];    // bit q: band q overlaps band r (filled at launch)
};

struct ProbeDesc {
  int kind, comp, i, j, k;
};

struct Probes {
  int n;
  const ProbeDesc* desc;
  float* series;  // [row][n]
  int row;
  int nfreq;
  const double* omega;
  double* dft;  // [n][nfreq][2]
  double dt;
  long long step;
};

// Poynting-flux faces (dtm_flux.hip).  Face normal to axis a on the H plane
// pa, tangential ranges [b0,b1] x [c0,c1] (cells, b = a+1, c = a+2 mod 3);
// `nodes` = (b1-b0+1)(c1-c0) + (b1-b0)(c1-c0+1) face nodes, whose E^n
// snapshot starts at snap0; area = d_b d_c; sign = -1 (low face) or +1.
constexpr int kMaxFluxBoxes = 8;
struct FluxFace {
  int a, pa, b0, b1, c0, c1;
  int nodes, snap0;
  double area, sign;
};
struct FluxSet {
  int nface;
  FluxFace face[6 * kMaxFluxBoxes];
  float* snap;     // face-averaged tangential E^n of every node
  double* series;  // [row][face]
  double* total;   // [face], running sum of flux * dt
};

struct Soft {
  int n;
  int comp[16];
  long long q[16];
  int ijk[16][3];  // cell of each source (the FP32 zone test)
  float val[16];  // dt * inv_eps * J(t)
};

// Launch status of the calling thread.  Every launcher calls note_launch
// right after its hipLaunchKernelGGL; the first failure (with the kernel's
// name) is kept until take_launch_error hands it to the C-ABI entry point,
// which then returns DTM_ERR_DEVICE.  A launcher that refuses a launch
// (a size it cannot index) records it with fail_launch.
void note_launch(const char* kernel);
void fail_launch(const char* kernel, const char* why);
// 1 and the message when a failure was recorded since the last call (cleared).
int take_launch_error(char* msg, int cap);

// Launchers (dtm_kernels.hip).  All enqueue on `st`; each sweep is one
// launch of the kernel variant with every boundary rule over `regs`.
// fold non-null: add a plane-wave TF/SF source along x inside the sweep,
// from this half-step's per-plane factor table (launch_tfsf_fold); ignored
// with an FP32 zone
// side: with an FP32 zone, the zone's column runs on side.st between the
// fork and join events (concurrently with the rest of the grid on st); null:
// both launches on st.
struct SideStream {
  hipStream_t st;
  hipEvent_t fork, join;
};
void launch_update_h(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                     const Region* regs, int nreg, hipStream_t st, const float* fold = nullptr,
                     const SideStream* side = nullptr);
void launch_update_e(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                     const Materials& mt, const Ade& ade, const Region* regs, int nreg, hipStream_t st,
                     const float* fold = nullptr, const SideStream* side = nullptr);
// floats per half-step of the fold factor table: (nx + 1) * 8
void launch_tfsf_fold(const Tfsf& tf, const double* tables, float* out, int nsteps, const Grid& g,
                      hipStream_t st);
// TF/SF corrections applied after a sweep (half = 0: H, 1: E).  The update is
// linear in the stencil sums, so the correction is additive, including its
// CPML psi and ADE J terms.
// mode 0: H post-correction, 1: E post-correction, 2: H pre-correction
// (applied to H^{n-1/2} and psi before a fused step: psi += c D / b).
// design-gradient adjoint (FP32 storage)
void launch_adj_snap(const Grid& g, const void* fields, const Materials& mt, float* out, hipStream_t st);
void launch_adj_monitor(const void* fields, const long long* q, int nmon, float* series, hipStream_t st);
void launch_adj_grad(const Grid& g, const void* fields, const Materials& mt, const float* e0, const float* e1,
                     double* acc, hipStream_t st);
void launch_adj_inject(const void* fields, const long long* q, const float* inv_eps, const double* F, int nmon,
                       double w, double dt, long long n, hipStream_t st);
// launch_adj_inject at step n, then launch_adj_grad (of step n-1), in one
// launch; only for monitors outside the design box.
void launch_adj_inject_grad(const Grid& g, const void* fields, const Materials& mt, const float* e0, const float* e1,
                            double* acc, const long long* q, const float* inv_eps, const double* F, int nmon,
                            double w, double dt, long long n, hipStream_t st);
void launch_tfsf_tables(const Tfsf& tf, double* out, long long n0, int nsteps, double dt, hipStream_t st);
void launch_tfsf(int store, int half, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                 const Materials& mt, const Ade& ade, const Tfsf& tf, const TfsfBands& bands, double t_inc,
                 hipStream_t st);
// Dispersive / gyrotropic cells over the ADE bounding box (after the E
// sweep): E^n read from fields_src, E^{n+1} written to fields_dst.
void launch_ade(int store, const Grid& g, const void* fields_src, const void* fields_dst, const Scales& sc,
                const Materials& mt, const Ade& ade, hipStream_t st);
// Fused time step (dtm_torre.hip): old fields/psi -> new fields/psi.
// `cold` is a device copy of the step's TorreCold (make_torre_cold).
// fold non-null: a plane-wave TF/SF source along x added inside the step,
// from the step's factor tables (the H half-step's, the E half-step's
// following it at +(nx + 1) * kFoldStride).
void launch_torre(int store, const Grid& g, const Scales& sc, const Materials& mt, const TorreCold* cold,
                  const Region* regs, int nreg, hipStream_t st, const float* fold = nullptr);
TorreCold make_torre_cold(const void* fields_old, const void* fields_new, int store, const Cpml& pm_old,
                          const Cpml& pm_new, const Materials& mt, const Ade& ade);
void launch_soft(int store, const Grid& g, const void* fields, const Scales& sc, const Soft& so,
                 hipStream_t st);
void launch_record(int store, const Grid& g, const void* fields, const Scales& sc, const Probes& pr,
                   hipStream_t st);
// Flux of the step n -> n+1: launch_flux_snap before it, launch_flux after it
// (dtm_flux.hip).
void launch_flux_snap(int store, const Grid& g, const void* fields, const Scales& sc, const FluxSet& fs,
                      hipStream_t st);
void launch_flux(int store, const Grid& g, const void* fields, const Scales& sc, const FluxSet& fs, int row,
                 double dt, hipStream_t st);
// 1/2 sum (eps E^2 + H^2) dv over a box in FP64: one partial per x plane of
// the box in part[0 .. i1-i0), then their fixed-order sum into *out.
void launch_energy_box(int store, const Grid& g, const void* fields, const Scales& sc, const Materials& mt,
                       const Region& bx, double* part, double dv, double* out, hipStream_t st);
// The same box energy with the partials in a static device array instead of
// allocated memory (dtm_flux.hip).  static_scratch_init looks the array up on
// device `dev`, which must be current (0, or -1 on failure);
// energy_box_static returns 0, -1 (device error) or -2 (more x planes than
// static_scratch_planes()).
int static_scratch_init(int dev);
int static_scratch_planes();
int energy_box_static(int dev, int store, const Grid& g, const void* fields, const Scales& sc, const Materials& mt,
                      const Region& bx, double dv, double* host_out, hipStream_t st);
void launch_to_float(int store, const void* src, float scale_inv, float* dst, long long n,
                     hipStream_t st);
void launch_from_float(int store, const float* src, float scale, void* dst, long long n,
                       hipStream_t st);
const char* kernel_variant_name(int store);

}  // namespace dtm
