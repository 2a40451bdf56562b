// dtm_kernels.hip — hand-written gfx950 kernels of the FDTD stepping path.
//
// One launch per half-step ("two-pass" sweep).  Each launch covers a set of
// boxes (regions): the interior box where every x/y stencil neighbour exists
// and only z can touch a boundary or a CPML slab, and up to four x/y edge
// slabs handled with full boundary logic — the zone split of the reference's
// Window::calcDtorres (window.hpp:126-145) without its host loop.
//
// Thread mapping: 256 threads = 64 (z, one wavefront) x 4 (y).  Blocks are
// remapped so that each of the 8 XCDs sweeps a contiguous run of (x,y) rows,
// keeping the x/y stencil neighbours of a block in that XCD's L2.
//
// Update equations (genAlg.py:182-301, constants defs.h:3-13), on the cell
// lattice of include/dtmaxwell4.h:
//   H^{n+1/2} = H^{n-1/2} - dt curl E^n          (body diamond, time 0)
//   E^{n+1}   = E^n + dt/eps curl H^{n+1/2}      (nose diamond, time 1/2)
// with the (2,4) difference D f = [27 (f_{+1/2} - f_{-1/2}) - (f_{+3/2} - f_{-3/2})]/(24 d).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include <algorithm>

#include "dtm_internal.h"

namespace dtm {

constexpr int BZ = 64;  // z extent of a block (one wavefront)
constexpr int BY = 4;   // y extent of a block
constexpr int NXCD = 8;

// ------------------------------------------------------- storage helpers
__device__ __forceinline__ float ld(const float* F, long long q, float) { return F[q]; }
__device__ __forceinline__ float ld(const __half* F, long long q, float inv) {
  return __half2float(F[q]) * inv;
}
__device__ __forceinline__ void st(float* F, long long q, float v, float) { F[q] = v; }
__device__ __forceinline__ void st(__half* F, long long q, float v, float s) {
  F[q] = __float2half_rn(v * s);
}
// one cell of component comp (Ex .. Hz = 0 .. 5) with the FP32 zone: xq is
// the cell's zone index (exact_idx) or -1; the half array always holds the
// rounded zone value, so readers that ignore the zone see FP16 precision
__device__ __forceinline__ float ldx(const float* F, long long q, float inv, const Exact&, int, int) {
  return ld(F, q, inv);
}
__device__ __forceinline__ float ldx(const __half* F, long long q, float inv, const Exact& xz, int comp, int xq) {
  return xq >= 0 ? xz.f[comp][xq] : ld(F, q, inv);
}
__device__ __forceinline__ void stx(float* F, long long q, float v, float s, const Exact&, int, int) {
  st(F, q, v, s);
}
__device__ __forceinline__ void stx(__half* F, long long q, float v, float s, const Exact& xz, int comp, int xq) {
  if (xq >= 0) xz.f[comp][xq] = v;
  st(F, q, v, s);
}
template <typename T>
__device__ __forceinline__ float ldn(const T* F, long long q, bool ok, float inv) {
  return ok ? ld(F, q, inv) : 0.0f;
}

// Neighbour index along one axis with the boundary rule: -1 => zero field.
__device__ __forceinline__ int nbr(int v, int o, int n, int per) {
  int w = v + o;
  if (w >= 0 && w < n) return w;
  if (!per) return -1;
  return w < 0 ? w + n : w - n;
}

// (2,4) difference with weights (+1, -27, +27, -1) on the samples at
// half-offsets (-3, -1, +1, +3) (genAlg.py:196-198, sign folded).
__device__ __forceinline__ float d24(float f0, float f1, float f2, float f3) {
  return 27.0f * (f2 - f1) - (f3 - f0);
}

// ------------------------------------------------------- incident fields
__device__ __forceinline__ void wave3(int waveform, double p, double w, double kEnv, double t0, double tw,
                      double& f0, double& f1, double& f2) {
  if (waveform == 0) {  // raised-cosine window (signal.hpp:71-77)
    const double kp = p * kEnv;
    if (kp < 0.0 || kp > 2.0 * M_PI) {
      f0 = f1 = f2 = 0.0;
      return;
    }
    double sp, cp, sk, ck;
    sincos(p, &sp, &cp);
    sincos(kp, &sk, &ck);
    f0 = sp * 0.5 * (1.0 - ck);
    f1 = 0.5 * cp * (1.0 - ck) + 0.5 * kEnv * sp * sk;
    f2 = -0.5 * sp * (1.0 - ck) + kEnv * cp * sk + 0.5 * kEnv * kEnv * sp * ck;
  } else {  // Gaussian-modulated sine, or the bare Gaussian (waveform 2, signal.hpp:55-57)
    const double a = 1.0 / (w * tw);
    const double u = (p / w - t0) / tw;
    const double g = exp(-u * u);
    const double g1 = -2.0 * u * a * g;
    const double g2 = (4.0 * u * u * a * a - 2.0 * a * a) * g;
    if (waveform == 2) {
      f0 = g;
      f1 = g1;
      f2 = g2;
      return;
    }
    double sp, cp;
    sincos(p, &sp, &cp);
    f0 = g * sp;
    f1 = g1 * sp + g * cp;
    f2 = g2 * sp + 2.0 * g1 * cp - g * sp;
  }
}

// Incident value of component comp (0..5) at half-coordinates h, time t.
// Dipole: signal.hpp:58-96; plane wave: signal.hpp:17-49.
__device__ __forceinline__ double incident(const Tfsf& s, int comp, int hx, int hy, int hz, double t) {
  const double x = hx * s.half_d[0], y = hy * s.half_d[1], z = hz * s.half_d[2];
  if (s.kind == 1) {
    double n0 = x - s.dip[0], n1 = y - s.dip[1], n2 = z - s.dip[2];
    const double R = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    const double c = s.cSrc, w = s.w;
    double f0, f1, f2;
    wave3(s.waveform, w * t - w / c * R, w, s.kEnv, s.t0, s.tw, f0, f1, f2);
    const double d = s.amp * f0, ds = s.amp * w * f1, dss = s.amp * w * w * f2;
    n0 /= R;
    n1 /= R;
    n2 /= R;
    // dipole along the unit axis u (theta, phi; theta = 0: the reference's z
    // dipole, signal.hpp:70-72): d = d u, so (dss x n) x n = dss ((u.n) n - u)
    const double u0 = s.st * s.cph, u1 = s.st * s.sph, u2 = s.ct;
    const double un = u0 * n0 + u1 * n1 + u2 * n2;
    if (comp < 3) {
      const double nc = comp == 0 ? n0 : (comp == 1 ? n1 : n2);
      const double uc = comp == 0 ? u0 : (comp == 1 ? u1 : u2);
      const double t2 = dss * (un * nc - uc);
      return c * (t2 / (c * c * R) + (3.0 * ds * un * nc - ds * uc) / (c * R * R) +
                  (3.0 * d * un * nc - d * uc) / (R * R * R));
    } else {
      // (v u) x n = v (u1 n2 - u2 n1, u2 n0 - u0 n2, u0 n1 - u1 n0)
      const int a = comp - 3;
      const double cz = a == 0 ? u1 * n2 - u2 * n1 : (a == 1 ? u2 * n0 - u0 * n2 : u0 * n1 - u1 * n0);
      return dss * cz / (c * c * R) + ds * cz / (c * R * R);
    }
  } else {
    const double phase = s.w * t - s.kx * (x - s.pos[0]) - s.ky * (y - s.pos[1]) - s.kz * (z - s.pos[2]);
    double f0, f1, f2;
    wave3(s.waveform, phase, s.w, s.kEnv, s.t0, s.tw, f0, f1, f2);
    const double sv = s.amp * f0;
    const double Ephi = s.Imp * sv * s.cps, Eth = s.Imp * sv * s.sps;
    const double Hth = -sv * s.cps, Hphi = sv * s.sps;
    switch (comp) {
      case 0: return -Ephi * s.sph + Eth * s.ct * s.cph;
      case 1: return Ephi * s.cph + Eth * s.ct * s.sph;
      case 2: return -Eth * s.st;
      case 3: return -Hphi * s.sph + Hth * s.ct * s.cph;
      case 4: return Hphi * s.cph + Hth * s.ct * s.sph;
      default: return -Hth * s.st;
    }
  }
}

__device__ __forceinline__ bool in_sf(const Tfsf& s, int hx, int hy, int hz) {
  const bool inside = hx >= s.hlo[0] && hx <= s.hhi[0] && hy >= s.hlo[1] && hy <= s.hhi[1] &&
                      hz >= s.hlo[2] && hz <= s.hhi[2];
  return s.sf_inside ? inside : !inside;
}

// Component living at half-coordinates (genAlg.py:24 typing).
__device__ __forceinline__ int comp_at(int hx, int hy, int hz) {
  const int o0 = hx & 1, o1 = hy & 1, o2 = hz & 1;
  if (o0 + o1 + o2 == 2) return o0 == 0 ? 0 : (o1 == 0 ? 1 : 2);
  return o0 ? 3 : (o1 ? 4 : 5);
}

// Sum over the 4 stencil neighbours along axis a of component-at-(hp) of the
// TF/SF corrections (genAlg.py:219-221) weighted by (+1,-27,+27,-1).
// TAB 1: plane wave with k along one axis, the incident value is the
// polarisation factor P[comp] times this half-step's waveform table entry;
// TAB 2: a dipole's point table (k_tfsf_points); TAB 0: evaluated here.
// The neighbours differ from the point only along a, so no face is crossed
// unless the point lies inside the box on the other two axes; then a
// neighbour crosses when its a-coordinate is on the other side of [hlo, hhi].
template <int TAB>
__device__ __forceinline__ float tfsf_dcorr(const Tfsf& s, int hx, int hy, int hz, int a, double t,
                                           const double* table) {
  const int hp[3] = {hx, hy, hz};
  const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
  if (hp[b1] < s.hlo[b1] || hp[b1] > s.hhi[b1] || hp[b2] < s.hlo[b2] || hp[b2] > s.hhi[b2]) return 0.f;
  const int h = hp[a], lo = s.hlo[a], hi = s.hhi[a];
  const bool pin = h >= lo && h <= hi;
  const int ho[4] = {-3, -1, 1, 3};
  const double wt[4] = {1.0, -27.0, 27.0, -1.0};
  // the four neighbours (odd offsets along a) hold one component, the same
  // for every cell of a band and component (wavefront-uniform)
  const int cq = __builtin_amdgcn_readfirstlane(comp_at(a == 0 ? hx + 1 : hx, a == 1 ? hy + 1 : hy, a == 2 ? hz + 1 : hz));
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int hq = h + ho[q];
    const bool qin = hq >= lo && hq <= hi;
    if (qin == pin) continue;
    const bool qsf = s.sf_inside ? qin : !qin;  // the neighbour is in the scattered-field region
    double v;
    if constexpr (TAB == 1) {
      const int ha = s.tab_axis == a ? hq : hp[s.tab_axis];
      v = s.P[cq] * table[min(max(ha - s.tab_h0, 0), s.tab_n - 1)];
    } else if constexpr (TAB == 2) {  // the launch's point table; clamped reads stay in it
      const int px = min(max((a == 0 ? hq : hx) - s.ptab_h0[0], 0), s.ptab_n[0] - 1);
      const int py = min(max((a == 1 ? hq : hy) - s.ptab_h0[1], 0), s.ptab_n[1] - 1);
      const int pz = min(max((a == 2 ? hq : hz) - s.ptab_h0[2], 0), s.ptab_n[2] - 1);
      v = s.ptab[(((long long)cq * s.ptab_n[0] + px) * s.ptab_n[1] + py) * s.ptab_n[2] + pz];
    } else {
      v = incident(s, cq, a == 0 ? hq : hx, a == 1 ? hq : hy, a == 2 ? hq : hz, t);
    }
    acc += wt[q] * (qsf ? v : -v);
  }
  return (float)acc;
}

// -------------------------------------------------------- CPML helpers
__device__ __forceinline__ int slab_index(int v, int n, int plo, int phi) {
  if (v < plo) return v;
  if (v >= n - phi) return plo + v - (n - phi);
  return -1;
}

__device__ __forceinline__ long long psi_index(const Grid& g, const Cpml& p, int a, int s, int i, int j,
                                               int k) {
  const int P = p.plo[a] + p.phi[a];
  if (a == 0) return ((long long)s * g.ny + j) * g.nz + k;
  if (a == 1) return ((long long)i * P + s) * g.nz + k;
  return ((long long)i * g.ny + j) * P + s;
}

// Split-field part update (genAlg.py:117-148): F = P1 + P2 with P2 stored,
// P1 <- k1a P1 + k2a ua and P2 <- k1b P2 + k2b ub, ua / ub the increments of
// the first and second curl term.  Outside every slab of the component
// (k2 = 1 on both axes) it is the plain update F + ua + ub.
__device__ __forceinline__ float split_step(float* P2, long long q, float f0, float k1a, float k2a, float ua,
                                            float k1b, float k2b, float ub) {
  if (k2a == 1.0f && k2b == 1.0f) return f0 + ua + ub;
  const float p2 = P2[q], p1 = f0 - p2;
  const float n1 = k1a * p1 + k2a * ua, n2 = k1b * p2 + k2b * ub;
  P2[q] = n2;
  return n1 + n2;
}

// 32-bit psi indices of the sweeps (psi arrays are far below 2^32 floats):
// z slab: base + slab index; x and y slabs: base + (k - k_lane0).
__device__ __forceinline__ unsigned psi_base_z(const Grid& g, const Cpml& p, int i, int j) {
  return ((unsigned)i * (unsigned)g.ny + (unsigned)j) * (unsigned)(p.plo[2] + p.phi[2]);
}
__device__ __forceinline__ unsigned psi_base_y(const Grid& g, const Cpml& p, int ys, int i, int k) {
  return ys < 0 ? 0u : ((unsigned)i * (unsigned)(p.plo[1] + p.phi[1]) + (unsigned)ys) * (unsigned)g.nz + (unsigned)k;
}
__device__ __forceinline__ unsigned psi_base_x(const Grid& g, const Cpml& p, int xs, int j, int k) {
  (void)p;
  return ((unsigned)xs * (unsigned)g.ny + (unsigned)j) * (unsigned)g.nz + (unsigned)k;
}

// psi <- b psi + c D ; D <- D + psi   (CPML, kappa=1, alpha=0)
__device__ __forceinline__ void cpml_apply(float* psi, long long pq, float b, float c, float& D) {
  float v = b * psi[pq] + c * D;
  psi[pq] = v;
  D += v;
}
// The same with the old psi value already loaded.
__device__ __forceinline__ void cpml_step(float* psi, long long pq, float old, float b, float c, float& D) {
  const float v = b * old + c * D;
  psi[pq] = v;
  D += v;
}

// ===================================================== marching sweeps
// A block owns a tile of 64 (z) x ROWS (y) columns and marches along x over
// a segment of planes (the tilted space-time march of the reference's torre,
// genAlg.py:420-499, with the rag shuffles of copyrags.cuh replaced by an
// LDS ring): the x stencil reads a 5-plane ring of the two transverse
// components, the y/z stencils a double-buffered LDS tile of the current
// plane with a 3-cell halo, and operands are prefetched PF planes ahead in
// registers.  One barrier per plane.
// y rows per block (ROWS): 8, or 16 with one cell per lane when the launch
// then fills the chip as well (see upd_h).  At 512x256x256 the FP32 E sweep
// ran 7% faster with 16 rows (the H sweep 1%); the paired FP16 sweeps ran 8%
// slower, and at 400x200x200 the 16-row launch left a fifth of the CUs idle.
constexpr int MLZ = 67;   // LDS row length: 64 + 3 halo columns

struct MarchSet {
  int n;
  Region r[6];
  int nbk[6], nbj[6], nbs[6];
  int start[7];
  int total, per, seg;
  int bz;  // z extent of a tile: BZ * (cells per lane)
  int by;  // y extent of a tile: the kernel's ROWS
};

static MarchSet make_march(const Region* regs, int nreg, int seg, int bz, int by) {
  MarchSet s{};
  s.seg = seg;
  s.bz = bz;
  s.by = by;
  int acc = 0, m = 0;
  for (int t = 0; t < nreg && m < 6; t++) {
    const Region& r = regs[t];
    if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
    s.r[m] = r;
    s.nbk[m] = (r.k1 - r.k0 + bz - 1) / bz;
    s.nbj[m] = (r.j1 - r.j0 + by - 1) / by;
    s.nbs[m] = (r.i1 - r.i0 + seg - 1) / seg;
    s.start[m] = acc;
    acc += s.nbk[m] * s.nbj[m] * s.nbs[m];
    m++;
  }
  s.n = m;
  s.start[m] = acc;
  s.total = acc;
  s.per = (acc + NXCD - 1) / NXCD;
  return s;
}

struct MarchTile {
  int r, i0, i1, j0, k0;
};

__device__ __forceinline__ bool march_locate(const MarchSet& ms, MarchTile& t) {
  const int b = blockIdx.x;
  const int L = (b % NXCD) * ms.per + b / NXCD;  // XCD-contiguous logical id
  if (L >= ms.total) return false;
  int r = 0;
  while (r + 1 < ms.n && L >= ms.start[r + 1]) r++;
  int loc = L - ms.start[r];
  const int bk = loc % ms.nbk[r];
  loc /= ms.nbk[r];
  const int bj = loc % ms.nbj[r];
  const int bs = loc / ms.nbj[r];
  const Region& R = ms.r[r];
  t.r = r;
  t.k0 = R.k0 + bk * ms.bz;
  t.j0 = R.j0 + bj * ms.by;
  t.i0 = R.i0 + bs * ms.seg;
  t.i1 = min(t.i0 + ms.seg, R.i1);
  return true;
}

// In-plane offset (j*sy + k) of a position under the boundary rule, or -1.
__device__ __forceinline__ long long plane_off(const Grid& g, int j, int k) {
  const int jj = nbr(j, 0, g.ny, g.per[1]), kk = nbr(k, 0, g.nz, g.per[2]);
  return (jj < 0 || kk < 0) ? -1LL : (long long)jj * g.sy + kk;
}
template <bool GEN>
__device__ __forceinline__ long long plane_base(const Grid& g, int i) {
  if (!GEN) return (long long)i * g.sx;
  const int ii = nbr(i, 0, g.nx, g.per[0]);
  return ii < 0 ? -1LL : (long long)ii * g.sx;
}
template <typename T>
__device__ __forceinline__ float ldp(const T* F, long long base, long long off, float inv) {
  return (base < 0 || off < 0) ? 0.0f : ld(F, base + off, inv);
}

// Barrier that waits only for this wave's LDS traffic, so global loads
// issued for the next plane stay in flight across it (__syncthreads would
// also drain every outstanding global load).  The "memory" clobber keeps the
// compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Raw storage element and its conversion: prefetched operands stay in the
// storage type until they are published to LDS, so no conversion (and no
// wait for the load) sits next to the load that produced them.
template <int STORE> struct Raw;
template <> struct Raw<0> { using T = float; };
template <> struct Raw<1> { using T = __half; };
__device__ __forceinline__ float cvt(float v, float) { return v; }
__device__ __forceinline__ float cvt(__half v, float inv) { return __builtin_fmaf(__half2float(v), inv, 0.0f); }
template <typename T> __device__ __forceinline__ T rzero();
template <> __device__ __forceinline__ float rzero<float>() { return 0.0f; }
template <> __device__ __forceinline__ __half rzero<__half>() { return __ushort_as_half((unsigned short)0); }
template <typename T>
__device__ __forceinline__ T ldr(const T* F, long long base, long long off) {
  return (base < 0 || off < 0) ? rzero<T>() : F[base + off];
}
// Address of a position under the boundary rule: outside the grid (base or
// off < 0) it is the array's zero pad, so the load itself is unconditional
// and its destination register never needs a merge (a merge would force a
// wait for the load before the next plane).  The index type I is 32-bit
// below 2^32 - kPad cells, which keeps the select out of 64-bit registers,
// and 64-bit in the wide variants (IxT).
template <bool W> struct IxT { using T = unsigned; };
template <> struct IxT<true> { using T = unsigned long long; };
template <typename T, typename I>
__device__ __forceinline__ const T* at(const T* F, long long base, long long off, I zidx) {
  const I idx = (base < 0 || off < 0) ? zidx : (I)(base + off);
  return F + idx;
}

// Ring slot of plane p in a 5-plane LDS ring (p >= -3).
__device__ __forceinline__ int ring5(int p) { return (p + 10) % 5; }

// ================================================================ H sweep
// V consecutive z cells per lane (V = 2 for FP16 storage on an even nz: the
// lane loads one __half2, so a prefetch register carries two cells, the
// same bytes in flight per register as FP32).  Raw word of V cells:
template <int STORE, int V> struct Vec;
template <> struct Vec<0, 1> { using P = float; };
template <> struct Vec<1, 1> { using P = __half; };
template <> struct Vec<1, 2> { using P = __half2; };
__device__ __forceinline__ void unpackv(float p, float, float* o) { o[0] = p; }
__device__ __forceinline__ void unpackv(__half p, float inv, float* o) {
  o[0] = __builtin_fmaf(__half2float(p), inv, 0.0f);
}
// The conversions are written as fma(x, scale, 0): one mixed-precision
// v_fma_mix (half operand, float result) per value instead of a convert and
// a multiply.  The scale is a power of two, so the product is exact and the
// result is the same as the convert-then-multiply.
__device__ __forceinline__ void unpackv(__half2 p, float inv, float* o) {
  const float2 f = __half22float2(p);
  o[0] = __builtin_fmaf(f.x, inv, 0.0f);
  o[1] = __builtin_fmaf(f.y, inv, 0.0f);
}
template <typename P> __device__ __forceinline__ P pzero();
template <> __device__ __forceinline__ float pzero<float>() { return 0.0f; }
template <> __device__ __forceinline__ __half pzero<__half>() { return __ushort_as_half((unsigned short)0); }
template <> __device__ __forceinline__ __half2 pzero<__half2>() { return __half2(pzero<__half>(), pzero<__half>()); }
// store V values at element q (q even for V = 2)
__device__ __forceinline__ void stv(float* F, long long q, const float* v, float) { F[q] = v[0]; }
__device__ __forceinline__ void stv(__half* F, long long q, const float* v, float s) {
  F[q] = __float2half_rn(__builtin_fmaf(v[0], s, 0.0f));
}
__device__ __forceinline__ void stv2(__half* F, long long q, const float* v, float s) {
  *reinterpret_cast<__half2*>(F + q) =
      __floats2half2_rn(__builtin_fmaf(v[0], s, 0.0f), __builtin_fmaf(v[1], s, 0.0f));
}
// word of V cells at a position under the boundary rule (see at())
template <typename P, typename T, typename I>
__device__ __forceinline__ P ldv(const T* F, long long base, long long off, I zidx) {
  const I idx = (base < 0 || off < 0) ? zidx : (I)(base + off);
  return *reinterpret_cast<const P*>(F + idx);
}

// Operands of one plane, loaded PF planes ahead: the Ex tile words of
// plane p (own, y-halo row) and the Ex z-halo value, the H read-modify-write
// words of plane p, and the Ey / Ez tile values of plane p+1 (the x stencil
// of H reaches one plane ahead).
template <typename P, typename T> struct PreH {
  P ex, exy, hx, hy, hz, ey, ez, ezy;
  T exz, eyz;
};

// GEN bit 0: x boundary rule and x CPML (per plane, block-uniform);
// bit 1: y CPML (per row).  The y and z boundary rules and the z CPML are
// always on: they are per-thread constants.
// SP: split-field absorbing layer instead of the CPML (one cell per lane).
// TF: a plane-wave TF/SF source along x folded in: the correction of
// genAlg.py:219-221 is added to the derivative it belongs to before the CPML
// and split-field updates, which are linear in it (the k_tfsf kernel applies
// the same terms after a sweep); fold: this half-step's factor table.
// W: 64-bit cell indices (grids of 2^32 - kPad cells and more).
template <int STORE, int GEN, int V, bool SP = false, int ROWS = 8, bool XZ = false, bool TF = false, bool W = false>
__global__ __launch_bounds__(BZ* ROWS, 4) void k_update_h(Grid g, Fields<STORE> f, Scales sc, Cpml pm, MarchSet ms,
                                                         const float* fold) {
  using T = typename Raw<STORE>::T;
  using P = typename Vec<STORE, V>::P;
  constexpr bool GX = GEN & 1, GY = GEN & 2;
  constexpr int BZV = BZ * V, LZ = BZV + 3, MBY = ROWS;
  __shared__ float sEx[2][MBY + 3][LZ];  // [row j-j0+2][col k-k0+2], plane i
  __shared__ float rEy[5][MBY][LZ];      // [row j-j0][col k-k0+2], ring of planes
  __shared__ float rEz[5][MBY + 3][BZV]; // [row j-j0+2][col k-k0], ring of planes
  __shared__ float sHo[3][MBY][BZV];     // this thread's H operands of plane i
  MarchTile t;
  if (!march_locate(ms, t)) return;
  const Region& R = ms.r[t.r];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int j = t.j0 + ty, k = t.k0 + V * tx;  // first of this lane's V cells
  const int c0 = V * tx;                       // its tile column
  const bool inreg = j < R.j1 && k < R.k1;     // V = 2 needs an even nz: cells pair up
  const long long own = plane_off(g, j, k);
  // halo positions: rows j0-2, j0-1, j0+MBY (waves ty<3); columns k0-2, k0-1, k0+BZV (lanes tx<3)
  const bool yw = ty < 3, zl = tx < 3;
  const int yrow = ty < 2 ? ty : MBY + 2;
  const long long yh = yw ? plane_off(g, t.j0 + (ty < 2 ? ty - 2 : MBY), k) : -1;
  const int zcol = tx < 2 ? tx : BZV + 2;
  const long long zh = zl ? plane_off(g, j, t.k0 + (tx < 2 ? tx - 2 : BZV)) : -1;
  const T* Ex = f.E[0];
  const T* Ey = f.E[1];
  const T* Ez = f.E[2];
  T* Hx = f.H[0];
  T* Hy = f.H[1];
  T* Hz = f.H[2];
  const float iex = sc.ie[0], iey = sc.ie[1], iez = sc.ie[2];
  const float ihx = sc.ih[0], ihy = sc.ih[1], ihz = sc.ih[2];
  using Ix = typename IxT<W>::T;
  const Ix zi = W ? (Ix)g.zidxw : (Ix)g.zidx;  // zero pad index

  const long long hq = inreg ? own : -1;  // H operands: owned cells only
  // FP32 zone (FP16 storage): a block whose tile (with its halo) meets the
  // zone replaces the LDS values of zone cells by their FP32 copies and
  // writes its new zone values there too
  static_assert(!XZ || STORE == 1, "the FP32 zone applies to FP16 storage");
  const Exact& xz = sc.xz;
  bool xtile = false;
  if constexpr (XZ)
    xtile = xz.on && (g.per[1] || g.per[2] ||
                      (xz.j0 < t.j0 + MBY + 3 && xz.j1 > t.j0 - 3 && xz.k0 < t.k0 + BZV + 3 && xz.k1 > t.k0 - 3));
  const int jo = j;
  const int jy = yw ? nbr(t.j0 + (ty < 2 ? ty - 2 : MBY), 0, g.ny, g.per[1]) : -1;
  const int kz = zl ? nbr(t.k0 + (tx < 2 ? tx - 2 : BZV), 0, g.nz, g.per[2]) : -1;
  auto ov = [&](float& dst, int comp, int pi, int pj, int pk) {
    // rows and lanes past the grid's edge hold the wrapped cells on a
    // periodic axis (plane_off), nothing otherwise
    pj = nbr(pj, 0, g.ny, g.per[1]);
    pk = nbr(pk, 0, g.nz, g.per[2]);
    const int xq = (pi < 0 || pj < 0 || pk < 0) ? -1 : exact_idx(xz, pi, pj, pk);
    if (xq >= 0) dst = xz.f[comp][xq];
  };
  auto fetch = [&](PreH<P, T>& S, int p) {
    const long long pb = plane_base<GX>(g, p);
    const long long pn = plane_base<GX>(g, p + 1);
    S.ex = ldv<P>(Ex, pb, own, zi);
    S.ey = ldv<P>(Ey, pn, own, zi);
    S.ez = ldv<P>(Ez, pn, own, zi);
    S.exy = ldv<P>(Ex, pb, yh, zi);
    S.ezy = ldv<P>(Ez, pn, yh, zi);
    S.exz = *at(Ex, pb, zh, zi);
    S.eyz = *at(Ey, pn, zh, zi);
    S.hx = ldv<P>((const T*)Hx, pb, hq, zi);
    S.hy = ldv<P>((const T*)Hy, pb, hq, zi);
    S.hz = ldv<P>((const T*)Hz, pb, hq, zi);
  };
  // Ey / Ez tile of plane p straight into its ring slot (prologue only)
  auto ring_direct = [&](int p) {
    const long long pb = plane_base<GX>(g, p);
    const int s = ring5(p);
    float v[V];
    unpackv(ldv<P>(Ey, pb, own, zi), iey, v);
#pragma unroll
    for (int q = 0; q < V; q++) rEy[s][ty][c0 + q + 2] = v[q];
    unpackv(ldv<P>(Ez, pb, own, zi), iez, v);
#pragma unroll
    for (int q = 0; q < V; q++) rEz[s][ty + 2][c0 + q] = v[q];
    if (yw) {
      unpackv(ldv<P>(Ez, pb, yh, zi), iez, v);
#pragma unroll
      for (int q = 0; q < V; q++) rEz[s][yrow][c0 + q] = v[q];
    }
    if (zl) rEy[s][ty][zcol] = cvt(ldr(Ey, pb, zh), iey);
    if constexpr (XZ) {
      if (xtile) {
        const int pp = GX ? nbr(p, 0, g.nx, g.per[0]) : p;
#pragma unroll
        for (int q = 0; q < V; q++) {
          ov(rEy[s][ty][c0 + q + 2], 1, pp, jo, k + q);
          ov(rEz[s][ty + 2][c0 + q], 2, pp, jo, k + q);
          if (yw) ov(rEz[s][yrow][c0 + q], 2, pp, jy, k + q);
        }
        if (zl) ov(rEy[s][ty][zcol], 1, pp, jo, kz);
      }
    }
  };

  const int ysl = (GY && pm.on) ? slab_index(j, g.ny, pm.plo[1], pm.phi[1]) : -1;
  const float rd0 = g.rd[0], rd1 = g.rd[1], rd2 = g.rd[2], dt = g.dt;
  // CPML state of this lane's cells (z and y are fixed per thread)
  int zs[V];
  float bz[V], cz[V];
#pragma unroll
  for (int q = 0; q < V; q++) {
    zs[q] = pm.on ? slab_index(k + q, g.nz, pm.plo[2], pm.phi[2]) : -1;
    bz[q] = zs[q] >= 0 ? pm.bH[2][k + q] : 0.f;
    cz[q] = zs[q] >= 0 ? pm.cH[2][k + q] : 0.f;
  }
  const float by = ysl >= 0 ? pm.bH[1][j] : 0.f, cy = ysl >= 0 ? pm.cH[1][j] : 0.f;
  // TF/SF fold: this lane's packed weights (zero away from the y and z faces)
  int tfw[V];
#pragma unroll
  for (int q = 0; q < V; q++) tfw[q] = 0;
  if constexpr (TF) {
    const int* hd = reinterpret_cast<const int*>(fold);
#pragma unroll
    for (int q = 0; q < V; q++) tfw[q] = fold_pack(0, j, k + q, hd);
  }

  // PF prefetch sets, each a plane ahead of the next: two planes in flight in
  // the launched variants, which carry the y-CPML (three would spill).
  constexpr int PF = (GY || V == 2) ? 2 : 3;
  PreH<P, T> S[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    S[u].ex = S[u].exy = S[u].hx = S[u].hy = S[u].hz = S[u].ey = S[u].ez = S[u].ezy = pzero<P>();
    S[u].exz = S[u].eyz = rzero<T>();
  }
  ring_direct(t.i0 - 2);
  ring_direct(t.i0 - 1);
  ring_direct(t.i0);
  // Prefetches are unconditional (a plane past the segment is clamped to
  // its last plane): with no branch around them the compiler's in-order
  // wait counts stay exact and a wait never reaches the other sets.
#pragma unroll
  for (int u = 0; u < PF; u++) fetch(S[u], min(t.i0 + u, t.i1 - 1));

  auto step = [&](int i, PreH<P, T>& Q) {
    const int b = i & 1;
    const int s0 = ring5(i - 2), s1 = ring5(i - 1), s2 = ring5(i), s3 = ring5(i + 1);
    // publish plane i (Ex) and plane i+1 (Ey, Ez); H operands to the own slots
    float v[V];
    unpackv(Q.ex, iex, v);
#pragma unroll
    for (int q = 0; q < V; q++) sEx[b][ty + 2][c0 + q + 2] = v[q];
    unpackv(Q.ey, iey, v);
#pragma unroll
    for (int q = 0; q < V; q++) rEy[s3][ty][c0 + q + 2] = v[q];
    unpackv(Q.ez, iez, v);
#pragma unroll
    for (int q = 0; q < V; q++) rEz[s3][ty + 2][c0 + q] = v[q];
    if (yw) {
      unpackv(Q.exy, iex, v);
#pragma unroll
      for (int q = 0; q < V; q++) sEx[b][yrow][c0 + q + 2] = v[q];
      unpackv(Q.ezy, iez, v);
#pragma unroll
      for (int q = 0; q < V; q++) rEz[s3][yrow][c0 + q] = v[q];
    }
    if (zl) {
      sEx[b][ty + 2][zcol] = cvt(Q.exz, iex);
      rEy[s3][ty][zcol] = cvt(Q.eyz, iey);
    }
    // the H operands go through this thread's own LDS slots: no register of
    // the prefetch set is live across its refetch
    unpackv(Q.hx, ihx, v);
#pragma unroll
    for (int q = 0; q < V; q++) sHo[0][ty][c0 + q] = v[q];
    unpackv(Q.hy, ihy, v);
#pragma unroll
    for (int q = 0; q < V; q++) sHo[1][ty][c0 + q] = v[q];
    unpackv(Q.hz, ihz, v);
#pragma unroll
    for (int q = 0; q < V; q++) sHo[2][ty][c0 + q] = v[q];
    if constexpr (XZ) {
      if (xtile && (g.per[0] || (i + 1 >= xz.i0 && i - 1 < xz.i1))) {  // Ex of plane i, Ey / Ez of i+1, own H of i
        const int ii = GX ? nbr(i, 0, g.nx, g.per[0]) : i;
        const int i1n = GX ? nbr(i + 1, 0, g.nx, g.per[0]) : i + 1;
#pragma unroll
        for (int q = 0; q < V; q++) {
          ov(sEx[b][ty + 2][c0 + q + 2], 0, ii, jo, k + q);
          ov(rEy[s3][ty][c0 + q + 2], 1, i1n, jo, k + q);
          ov(rEz[s3][ty + 2][c0 + q], 2, i1n, jo, k + q);
          ov(sHo[0][ty][c0 + q], 3, ii, jo, k + q);
          ov(sHo[1][ty][c0 + q], 4, ii, jo, k + q);
          ov(sHo[2][ty][c0 + q], 5, ii, jo, k + q);
          if (yw) {
            ov(sEx[b][yrow][c0 + q + 2], 0, ii, jy, k + q);
            ov(rEz[s3][yrow][c0 + q], 2, i1n, jy, k + q);
          }
        }
        if (zl) {
          ov(sEx[b][ty + 2][zcol], 0, ii, jo, kz);
          ov(rEy[s3][ty][zcol], 1, i1n, jo, kz);
        }
      }
    }
    // CPML psi of plane i, loaded before the prefetch so that waiting for
    // psi does not wait for it (loads complete in issue order)
    // (z and y slabs; the x slab is a few planes, its psi is read in place)
    float pz0[V], pz1[V], py0[V], py1[V];
    const unsigned zb = psi_base_z(g, pm, i, j), yb = psi_base_y(g, pm, ysl, i, k);
#pragma unroll
    for (int q = 0; q < V; q++) {
      pz0[q] = pz1[q] = py0[q] = py1[q] = 0.f;
      if (!SP && inreg && zs[q] >= 0) {
        pz0[q] = pm.psiH[2][0][zb + zs[q]];
        pz1[q] = pm.psiH[2][1][zb + zs[q]];
      }
      if (!SP && GY && inreg && ysl >= 0) {
        py0[q] = pm.psiH[1][0][yb + q];
        py1[q] = pm.psiH[1][1][yb + q];
      }
    }
    const int xsl = (GX && pm.on) ? slab_index(i, g.nx, pm.plo[0], pm.phi[0]) : -1;
    // this plane's fold factors, loaded before the barrier (whose wait covers
    // them) rather than inside the band branch
    float fu[6];
#pragma unroll
    for (int t6 = 0; t6 < 6; t6++) fu[t6] = 0.f;
    if constexpr (TF) {
#pragma unroll
      for (int t6 = 0; t6 < 6; t6++) fu[t6] = fold[kFoldStride * (1 + i) + t6];
    }
    fetch(Q, min(i + PF, t.i1 - 1));
    lds_barrier();
    if (inreg) {
      float nx_[V], ny_[V], nz_[V];
#pragma unroll
      for (int q = 0; q < V; q++) {
        const int c = c0 + q;
        const float hx0 = sHo[0][ty][c], hy0 = sHo[1][ty][c], hz0 = sHo[2][ty][c];
        const float exo = sEx[b][ty + 2][c + 2];
        float dyEz = rd1 * d24(rEz[s2][ty][c], rEz[s2][ty + 1][c], rEz[s2][ty + 2][c], rEz[s2][ty + 3][c]);
        float dyEx = rd1 * d24(sEx[b][ty][c + 2], sEx[b][ty + 1][c + 2], exo, sEx[b][ty + 3][c + 2]);
        float dzEy = rd2 * d24(rEy[s2][ty][c], rEy[s2][ty][c + 1], rEy[s2][ty][c + 2], rEy[s2][ty][c + 3]);
        float dzEx = rd2 * d24(sEx[b][ty + 2][c], sEx[b][ty + 2][c + 1], exo, sEx[b][ty + 2][c + 3]);
        float dxEz = rd0 * d24(rEz[s0][ty + 2][c], rEz[s1][ty + 2][c], rEz[s2][ty + 2][c], rEz[s3][ty + 2][c]);
        float dxEy = rd0 * d24(rEy[s0][ty][c + 2], rEy[s1][ty][c + 2], rEy[s2][ty][c + 2], rEy[s3][ty][c + 2]);
        if constexpr (TF) {
          if (tfw[q]) {  // a lane next to a y or z face (or inside the box on both)
            const int w = tfw[q];
            dyEz += fold_f(w, 0) * fu[0];   // Hx, y faces
            dzEy += fold_f(w, 6) * fu[1];   // Hx, z faces
            dzEx += fold_f(w, 12) * fu[2];  // Hy, z faces
            dxEz += fold_b(w, 24) * fu[3];  // Hy, x faces
            dxEy += fold_b(w, 25) * fu[4];  // Hz, x faces
            dyEx += fold_f(w, 18) * fu[5];  // Hz, y faces
          }
        }
        if constexpr (SP) {
          // split parts, (k1, k2) = (1, 1) off the slab: Hx (y, z), Hy (z, x), Hz (x, y)
          const bool sx = GX && xsl >= 0;
          const float k1x = sx ? pm.bH[0][i] : 1.f, k2x = sx ? pm.cH[0][i] : 1.f;
          const float k1y = (GY && ysl >= 0) ? by : 1.f, k2y = (GY && ysl >= 0) ? cy : 1.f;
          const float k1z = zs[q] >= 0 ? bz[q] : 1.f, k2z = zs[q] >= 0 ? cz[q] : 1.f;
          const long long qs = plane_base<GX>(g, i) + own + q;
          nx_[q] = split_step(pm.p2H[0], qs, hx0, k1y, k2y, -dt * dyEz, k1z, k2z, dt * dzEy);
          ny_[q] = split_step(pm.p2H[1], qs, hy0, k1z, k2z, -dt * dzEx, k1x, k2x, dt * dxEz);
          nz_[q] = split_step(pm.p2H[2], qs, hz0, k1x, k2x, -dt * dxEy, k1y, k2y, dt * dyEx);
        } else {
          if (zs[q] >= 0) {  // z axis: slot 0 = Hx (D_z Ey), slot 1 = Hy (D_z Ex)
            cpml_step(pm.psiH[2][0], zb + zs[q], pz0[q], bz[q], cz[q], dzEy);
            cpml_step(pm.psiH[2][1], zb + zs[q], pz1[q], bz[q], cz[q], dzEx);
          }
          if (GX && xsl >= 0) {  // x axis: slot 0 = Hy (D_x Ez), slot 1 = Hz (D_x Ey)
            const unsigned qx = psi_base_x(g, pm, xsl, j, k) + q;
            cpml_apply(pm.psiH[0][0], qx, pm.bH[0][i], pm.cH[0][i], dxEz);
            cpml_apply(pm.psiH[0][1], qx, pm.bH[0][i], pm.cH[0][i], dxEy);
          }
          if (GY && ysl >= 0) {  // y axis: slot 0 = Hz (D_y Ex), slot 1 = Hx (D_y Ez)
            cpml_step(pm.psiH[1][0], yb + q, py0[q], by, cy, dyEx);
            cpml_step(pm.psiH[1][1], yb + q, py1[q], by, cy, dyEz);
          }
          nx_[q] = hx0 - dt * (dyEz - dzEy);
          ny_[q] = hy0 - dt * (dzEx - dxEz);
          nz_[q] = hz0 - dt * (dxEy - dyEx);
        }
      }
      const long long qq = plane_base<GX>(g, i) + own;
      if constexpr (XZ) {
        if (xtile) {  // new H of zone cells to the FP32 copy
#pragma unroll
          for (int q = 0; q < V; q++) {
            const int xq = exact_idx(xz, i, j, k + q);
            if (xq >= 0) {
              xz.f[3][xq] = nx_[q];
              xz.f[4][xq] = ny_[q];
              xz.f[5][xq] = nz_[q];
            }
          }
        }
      }
      if constexpr (V == 2) {
        stv2(Hx, qq, nx_, sc.sh[0]);
        stv2(Hy, qq, ny_, sc.sh[1]);
        stv2(Hz, qq, nz_, sc.sh[2]);
      } else {
        stv(Hx, qq, nx_, sc.sh[0]);
        stv(Hy, qq, ny_, sc.sh[1]);
        stv(Hz, qq, nz_, sc.sh[2]);
      }
    }
  };
  int i = t.i0;
  for (; i + PF - 1 < t.i1; i += PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) step(i + u, S[u]);
  }
#pragma unroll
  for (int u = 0; u < PF - 1; u++)
    if (i + u < t.i1) step(i + u, S[u]);
}

// ============================================================ ADE cells
// Dispersive / gyrotropic cells: countDisp (dispersion.cpp:85-99) and the
// magnetised-plasma current, from the curl H the E sweep stored.
template <int STORE>
__global__ __launch_bounds__(256) void k_ade(Grid g, Fields<STORE> f, Fields<STORE> fd, Scales sc, Materials mt,
                                             Ade ade) {
  const long long nb = (long long)ade.ni * ade.nj * ade.nk;
  for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nb; b += (long long)gridDim.x * blockDim.x) {
    const int k = ade.k0 + (int)(b % ade.nk);
    const int j = ade.j0 + (int)((b / ade.nk) % ade.nj);
    const int i = ade.i0 + (int)(b / ((long long)ade.nk * ade.nj));
    const long long q = i * g.sx + j * g.sy + k;
    const ClassEntry ce = mt.table[mt.cls ? mt.cls[q] : 0];
    if (!ce.special) continue;
    if (mt.d_on) {
      const int di = i - mt.d_i0, dj = j - mt.d_j0, dk = k - mt.d_k0;
      if (di >= 0 && di < mt.d_ni && dj >= 0 && dj < mt.d_nj && dk >= 0 && dk < mt.d_nk) continue;
    }
    float curl[3] = {ade.curl[0][b], ade.curl[1][b], ade.curl[2][b]};
    const int xq = STORE == 1 ? exact_idx(sc.xz, i, j, k) : -1;
    float En[3];
    for (int c = 0; c < 3; c++) En[c] = ldx(f.E[c], q, sc.ie[c], sc.xz, c, xq);
    if ((ce.special & 8) && ade.gyro_on) {
      const GyroMat& gm = ade.gmats[ce.mat[0]];
      const float J0[3] = {ade.Jg[0][b], ade.Jg[1][b], ade.Jg[2][b]};
      for (int r = 0; r < 3; r++) {
        float v = 0.f;
        for (int t = 0; t < 3; t++) v += gm.A[3 * r + t] * J0[t] + gm.B[3 * r + t] * En[t];
        ade.Jg[r][b] = v;
        curl[r] -= v;
      }
    }
    for (int c = 0; c < 3; c++) {
      if (ce.special & (1 << c)) {
        const AdeMat& am = ade.mats[ce.mat[c]];
        const float Ep = En[c];
        const float Em = ade.Em[c][b];
        const float rotdt = g.dt * curl[c];
        float sum = 0.f;
        for (int p = 0; p < am.np; p++) sum += am.kEJ[p][0] * ade.Jp[c][p][b] + am.kEJ[p][1] * ade.Jm[c][p][b];
        const float Enew = am.kE[0] * Em + am.kE[1] * Ep + am.kE[2] * rotdt - sum;
        for (int p = 0; p < am.np; p++) {
          const float jp = ade.Jp[c][p][b], jm = ade.Jm[c][p][b];
          const float jn = am.kJ[p][0] * jp + am.kJ[p][1] * jm + am.kJ[p][2] * Enew + am.kJ[p][3] * Em +
                           am.kJ[p][4] * Ep;
          ade.Jm[c][p][b] = jp;
          ade.Jp[c][p][b] = jn;
        }
        ade.Em[c][b] = Ep;
        stx(fd.E[c], q, Enew, sc.se[c], sc.xz, c, xq);
      } else {
        stx(fd.E[c], q, En[c] + g.dt * ce.inv_eps[c] * curl[c], sc.se[c], sc.xz, c, xq);
      }
    }
  }
}

// ====================================================== TF/SF correction
// For a cell in the face band of axis a, the two components with a
// derivative along a get the TF/SF term of genAlg.py:219-221 added to their
// stencil sum.  The sweep is linear in that sum, so the correction is
//   H_c -= dt * s * (1 + c_pml) * D,   psi += c_pml * D      (H half)
//   E_c += dt * ie * s * (1 + c_pml) * D, psi += c_pml * D   (E half)
// with s = +1 for the first curl term and -1 for the second; a dispersive
// component gets C3 * s * (1+c_pml) * D and its J^{n+1} the matching
// kJ2 * dE (dispersion.cpp:84-99 is linear in the curl).
// MODE 2 (H pre-correction before a fused step, on H^{n-1/2}): the fused
// step's own psi update then carries c*D, so psi gets c*D/b and H gets
// -dt*s*D only.
// One thread per cell of the union of the face bands: the cell belongs to
// the first band that contains it, and that thread applies the terms of
// every axis whose band holds the cell, so cells where bands of different
// axes overlap (box edges and corners) are written by one thread only.
template <int STORE, int MODE, int TAB>
__global__ __launch_bounds__(256) void k_tfsf(Grid g, Fields<STORE> f, Scales sc, Cpml pm, Materials mt, Ade ade,
                                              Tfsf tf, TfsfBands bands, double tinc) {
  constexpr int HALF = MODE == 1 ? 1 : 0;
  // band and position from the thread index: 32-bit arithmetic (band cells
  // number far fewer than 2^31; 64-bit division would dominate the kernel).
  // Each band starts at a multiple of 64 threads, so a wavefront lies in one
  // band: the band index is found from the wavefront's first thread and the
  // band tables are read with scalar loads.
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned g0 = __builtin_amdgcn_readfirstlane(gid);
  int r = 0;
  while (r < bands.n && g0 >= bands.start[r + 1]) r++;
  if (r >= bands.n) return;
  const Region& R = bands.r[r];
  const unsigned nk = (unsigned)(R.k1 - R.k0), nj = (unsigned)(R.j1 - R.j0);
  unsigned loc = gid - bands.start[r];
  if (loc >= nk * nj * (unsigned)(R.i1 - R.i0)) return;  // the band's padding
  const unsigned lk = loc % nk;
  loc /= nk;
  const int k = R.k0 + (int)lk;
  const int j = R.j0 + (int)(loc % nj);
  const int i = R.i0 + (int)(loc / nj);
  auto holds = [&](int q) {  // cell in band q (only bands that overlap band r can hold it)
    const Region& Q = bands.r[q];
    return ((bands.ovl[r] >> q) & 1u) && i >= Q.i0 && i < Q.i1 && j >= Q.j0 && j < Q.j1 && k >= Q.k0 &&
           k < Q.k1;
  };
  for (int r2 = 0; r2 < r; r2++)
    if (holds(r2)) return;  // owned by an earlier band
  const long long q = i * g.sx + j * g.sy + k;
  const int vpos[3] = {i, j, k};
  static constexpr int off[6][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  float dc[3] = {0.f, 0.f, 0.f};  // curl correction per component, all axes
  float dc2[3] = {0.f, 0.f, 0.f}; // its part on the stored split-field part (second curl axis)
  for (int rb = r; rb < bands.n; rb++) {
    if (rb != r && !holds(rb)) continue;
    const int a = bands.axis[rb];
    // components with a derivative along a: c1 = (a+2)%3 (first curl term),
    // c2 = (a+1)%3 (second curl term); psi slots 1 and 0 of axis a.
    for (int t = 0; t < 2; t++) {
      const int c = t == 0 ? (a + 2) % 3 : (a + 1) % 3;
      const float sgn = t == 0 ? 1.0f : -1.0f;
      const int slot = t == 0 ? 1 : 0;
      const int comp = HALF == 0 ? 3 + c : c;
      const int hx = 2 * i + off[comp][0], hy = 2 * j + off[comp][1], hz = 2 * k + off[comp][2];
      const float D = g.rd[a] * tfsf_dcorr<TAB>(tf, hx, hy, hz, a, tinc, tf.table);
      if (D == 0.0f) continue;
      float cp = 0.0f;
      if (pm.on && pm.split) {  // split-field layer: the part of axis a gets k2 * D
        auto k2_of = [&](int ax) {  // 1 off the slab of axis ax
          const int n = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
          if (slab_index(vpos[ax], n, pm.plo[ax], pm.phi[ax]) < 0) return 1.0f;
          return HALF == 0 ? pm.cH[ax][vpos[ax]] : pm.cE[ax][vpos[ax]];
        };
        const float k2 = k2_of(a);
        // the cell keeps split parts when either curl axis has k2 != 1; the
        // second curl axis (t = 1) owns the stored part
        if (k2 != 1.0f || k2_of(t == 0 ? (c + 2) % 3 : (c + 1) % 3) != 1.0f) {
          dc[c] += sgn * k2 * D;
          if (t == 1) dc2[c] += sgn * k2 * D;
          continue;
        }
      } else if (pm.on) {
        const int n = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
        const int si = slab_index(vpos[a], n, pm.plo[a], pm.phi[a]);
        if (si >= 0) {
          cp = HALF == 0 ? pm.cH[a][vpos[a]] : pm.cE[a][vpos[a]];
          float* psi = HALF == 0 ? pm.psiH[a][slot] : pm.psiE[a][slot];
          const long long pq = psi_index(g, pm, a, si, i, j, k);
          if (MODE == 2) {
            psi[pq] += cp * D / pm.bH[a][vpos[a]];
            cp = 0.0f;
          } else {
            psi[pq] += cp * D;
          }
        }
      }
      dc[c] += sgn * (1.0f + cp) * D;
    }
  }
  const int xq = STORE == 1 ? exact_idx(sc.xz, i, j, k) : -1;
  if (HALF == 0) {
    for (int c = 0; c < 3; c++) {
      if (dc[c] != 0.0f)
        stx(f.H[c], q, ldx(f.H[c], q, sc.ih[c], sc.xz, 3 + c, xq) - g.dt * dc[c], sc.sh[c], sc.xz, 3 + c, xq);
      if (dc2[c] != 0.0f) pm.p2H[c][q] -= g.dt * dc2[c];
    }
    return;
  }
  const ClassEntry ce = mt.table[mt.cls ? mt.cls[q] : 0];
  bool over = false;
  float dv = 0.f;
  if (mt.d_on) {
    const int di = i - mt.d_i0, dj = j - mt.d_j0, dk = k - mt.d_k0;
    if (di >= 0 && di < mt.d_ni && dj >= 0 && dj < mt.d_nj && dk >= 0 && dk < mt.d_nk) {
      dv = mt.d_inv_eps[((long long)di * mt.d_nj + dj) * mt.d_nk + dk];
      over = true;
    }
  }
  for (int c = 0; c < 3; c++) {
    if (dc[c] == 0.0f) continue;
    if (!over && (ce.special & (1 << c)) && ade.on) {
      const long long b = (long long)(i - ade.i0) * ade.bsx + (long long)(j - ade.j0) * ade.bsy + (k - ade.k0);
      const AdeMat& am = ade.mats[ce.mat[c]];
      const float dE = am.kE[2] * g.dt * dc[c];
      stx(f.E[c], q, ldx(f.E[c], q, sc.ie[c], sc.xz, c, xq) + dE, sc.se[c], sc.xz, c, xq);
      for (int p = 0; p < am.np; p++) ade.Jp[c][p][b] += am.kJ[p][2] * dE;
    } else {
      const float ie = over ? dv : ce.inv_eps[c];
      stx(f.E[c], q, ldx(f.E[c], q, sc.ie[c], sc.xz, c, xq) + g.dt * ie * dc[c], sc.se[c], sc.xz, c, xq);
      if (dc2[c] != 0.0f) pm.p2E[c][q] += g.dt * ie * dc2[c];
    }
  }
}

// Plane-wave waveform along the propagation axis, one value per
// half-coordinate: with k parallel to axis a the incident phase depends on
// the half-coordinate along a only (signal.hpp:17-26).  One launch fills the
// tables of nsteps steps, both half-steps: out[(s*2 + half)*tab_n + h] at
// t = (n0+s) dt (H half, incident E) and (n0+s+1/2) dt (E half, incident H).
__global__ __launch_bounds__(256) void k_tfsf_table(Tfsf tf, double* out, long long n0, int nsteps, double dt) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nsteps * 2 * tf.tab_n) return;
  const int h = (int)(idx % tf.tab_n);
  const int half = (int)((idx / tf.tab_n) % 2);
  const long long n = n0 + idx / (2LL * tf.tab_n);
  const double t = half ? ((double)n + 0.5) * dt : (double)n * dt;
  const int a = tf.tab_axis;
  const double ka = a == 0 ? tf.kx : (a == 1 ? tf.ky : tf.kz);
  const double xa = (h + tf.tab_h0) * tf.half_d[a];
  double f0, f1, f2;
  wave3(tf.waveform, tf.w * t - ka * (xa - tf.pos[a]), tf.w, tf.kEnv, tf.t0, tf.tw, f0, f1, f2);
  out[idx] = tf.amp * f0;
}

// ================================================================ E sweep
// Operands of one plane, loaded PF planes ahead: the Hx tile words of
// plane p (own, y-halo row) and the Hx z-halo value, the E read-modify-write
// words and class bytes of plane p, and the Hy / Hz tile values of plane p+2
// (the x stencil of E reaches two planes ahead).
template <typename P, typename T, typename CL> struct PreE {
  P hx, hxy, hy, hz, hzy, ex, ey, ez;
  T hxz, hyz;
  CL cls;
};
template <int V> struct ClsWord;
template <> struct ClsWord<1> { using T = uint8_t; };
template <> struct ClsWord<2> { using T = unsigned short; };

// MAT: the grid may hold dispersive / gyrotropic cells or a design box; the
// variants without (MAT false) carry none of their branches and pointers.
template <int STORE, int GEN, int V, bool SP = false, int ROWS = 8, bool XZ = false, bool TF = false, bool MAT = true,
          bool W = false>
__global__ __launch_bounds__(BZ* ROWS, 4) void k_update_e(Grid g, Fields<STORE> f, Scales sc, Cpml pm, Materials mt,
                                                         Ade ade, MarchSet ms, const float* fold) {
  using T = typename Raw<STORE>::T;
  using P = typename Vec<STORE, V>::P;
  using CL = typename ClsWord<V>::T;
  constexpr bool GX = GEN & 1, GY = GEN & 2;
  constexpr bool DES = MAT && V == 1;  // design-box support (its launches use one cell per lane)
  constexpr int BZV = BZ * V, LZ = BZV + 3, MBY = ROWS;
  __shared__ ClassEntry tbl[kMaxClasses];
  __shared__ float sHx[2][MBY + 3][LZ];  // [row j-j0+1][col k-k0+1], plane i
  __shared__ float rHy[5][MBY][LZ];      // [row j-j0][col k-k0+1], ring of planes
  __shared__ float rHz[5][MBY + 3][BZV]; // [row j-j0+1][col k-k0], ring of planes
  __shared__ float sEo[3][MBY][BZV];     // this thread's E operands of plane i
  __shared__ uint8_t sCl[MBY][BZV];      // and its class bytes
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int q = tx + ty * BZ; q < mt.nclass; q += BZ * MBY) tbl[q] = mt.table[q];
  MarchTile t;
  if (!march_locate(ms, t)) return;  // block-uniform
  const Region& R = ms.r[t.r];
  const int j = t.j0 + ty, k = t.k0 + V * tx;
  const int c0 = V * tx;
  const bool inreg = j < R.j1 && k < R.k1;
  const long long own = plane_off(g, j, k);
  // halo rows j0-1, j0+MBY, j0+MBY+1 (waves ty<3); columns k0-1, k0+BZV, k0+BZV+1 (lanes tx<3)
  const bool yw = ty < 3, zl = tx < 3;
  const int yrow = ty == 0 ? 0 : MBY + ty;
  const long long yh = yw ? plane_off(g, t.j0 + (ty == 0 ? -1 : MBY + ty - 1), k) : -1;
  const int zcol = tx == 0 ? 0 : BZV + tx;
  const long long zh = zl ? plane_off(g, j, t.k0 + (tx == 0 ? -1 : BZV + tx - 1)) : -1;
  const T* Hx = f.H[0];
  const T* Hy = f.H[1];
  const T* Hz = f.H[2];
  T* Ex = f.E[0];
  T* Ey = f.E[1];
  T* Ez = f.E[2];
  const float ihx = sc.ih[0], ihy = sc.ih[1], ihz = sc.ih[2];
  const float iex = sc.ie[0], iey = sc.ie[1], iez = sc.ie[2];
  using Ix = typename IxT<W>::T;
  const Ix zi = W ? (Ix)g.zidxw : (Ix)g.zidx;  // zero pad index

  const long long eq = inreg ? own : -1;  // E operands: owned cells only
  // FP32 zone (FP16 storage), as in the H sweep
  static_assert(!XZ || STORE == 1, "the FP32 zone applies to FP16 storage");
  const Exact& xz = sc.xz;
  bool xtile = false;
  if constexpr (XZ)
    xtile = xz.on && (g.per[1] || g.per[2] ||
                      (xz.j0 < t.j0 + MBY + 3 && xz.j1 > t.j0 - 3 && xz.k0 < t.k0 + BZV + 3 && xz.k1 > t.k0 - 3));
  const int jo = j;
  const int jy = yw ? nbr(t.j0 + (ty == 0 ? -1 : MBY + ty - 1), 0, g.ny, g.per[1]) : -1;
  const int kz = zl ? nbr(t.k0 + (tx == 0 ? -1 : BZV + tx - 1), 0, g.nz, g.per[2]) : -1;
  auto ov = [&](float& dst, int comp, int pi, int pj, int pk) {
    // rows and lanes past the grid's edge hold the wrapped cells on a
    // periodic axis (plane_off), nothing otherwise
    pj = nbr(pj, 0, g.ny, g.per[1]);
    pk = nbr(pk, 0, g.nz, g.per[2]);
    const int xq = (pi < 0 || pj < 0 || pk < 0) ? -1 : exact_idx(xz, pi, pj, pk);
    if (xq >= 0) dst = xz.f[comp][xq];
  };
  // with no class array every cell is class 0: read the zero page
  const uint8_t* clsp = mt.cls ? mt.cls : static_cast<const uint8_t*>(g.zero);
  const long long cq = mt.cls ? eq : 0;
  const Ix czidx = mt.cls ? zi : (Ix)0;
  auto fetch = [&](PreE<P, T, CL>& S, int p) {
    const long long pb = plane_base<GX>(g, p);
    const long long pn = plane_base<GX>(g, p + 2);
    S.hx = ldv<P>(Hx, pb, own, zi);
    S.hy = ldv<P>(Hy, pn, own, zi);
    S.hz = ldv<P>(Hz, pn, own, zi);
    S.hxy = ldv<P>(Hx, pb, yh, zi);
    S.hzy = ldv<P>(Hz, pn, yh, zi);
    S.hxz = *at(Hx, pb, zh, zi);
    S.hyz = *at(Hy, pn, zh, zi);
    S.ex = ldv<P>((const T*)Ex, pb, eq, zi);
    S.ey = ldv<P>((const T*)Ey, pb, eq, zi);
    S.ez = ldv<P>((const T*)Ez, pb, eq, zi);
    // planes without a class other than 0 read the class array's zero pad
    // (one cached line) instead of a byte per cell
    const int pw = GX ? nbr(p, 0, g.nx, g.per[0]) : p;
    S.cls = ldv<CL>(clsp, mt.cls ? (pw >= mt.cx0 && pw < mt.cx1 ? pb : -1) : 0, cq, czidx);
  };
  auto ring_direct = [&](int p) {  // Hy / Hz tile of plane p into its slot
    const long long pb = plane_base<GX>(g, p);
    const int s = ring5(p);
    float v[V];
    unpackv(ldv<P>(Hy, pb, own, zi), ihy, v);
#pragma unroll
    for (int q = 0; q < V; q++) rHy[s][ty][c0 + q + 1] = v[q];
    unpackv(ldv<P>(Hz, pb, own, zi), ihz, v);
#pragma unroll
    for (int q = 0; q < V; q++) rHz[s][ty + 1][c0 + q] = v[q];
    if (yw) {
      unpackv(ldv<P>(Hz, pb, yh, zi), ihz, v);
#pragma unroll
      for (int q = 0; q < V; q++) rHz[s][yrow][c0 + q] = v[q];
    }
    if (zl) rHy[s][ty][zcol] = cvt(ldr(Hy, pb, zh), ihy);
    if constexpr (XZ) {
      if (xtile) {
        const int pp = GX ? nbr(p, 0, g.nx, g.per[0]) : p;
#pragma unroll
        for (int q = 0; q < V; q++) {
          ov(rHy[s][ty][c0 + q + 1], 4, pp, jo, k + q);
          ov(rHz[s][ty + 1][c0 + q], 5, pp, jo, k + q);
          if (yw) ov(rHz[s][yrow][c0 + q], 5, pp, jy, k + q);
        }
        if (zl) ov(rHy[s][ty][zcol], 4, pp, jo, kz);
      }
    }
  };

  const int ysl = (GY && pm.on) ? slab_index(j, g.ny, pm.plo[1], pm.phi[1]) : -1;
  bool design_col[V];
  int zs[V];
  float bz[V], cz[V];
#pragma unroll
  for (int q = 0; q < V; q++) {
    design_col[q] = DES && mt.d_on && j >= mt.d_j0 && j < mt.d_j0 + mt.d_nj && k + q >= mt.d_k0 &&
                    k + q < mt.d_k0 + mt.d_nk;
    zs[q] = pm.on ? slab_index(k + q, g.nz, pm.plo[2], pm.phi[2]) : -1;
    bz[q] = zs[q] >= 0 ? pm.bE[2][k + q] : 0.f;
    cz[q] = zs[q] >= 0 ? pm.cE[2][k + q] : 0.f;
  }
  const float rd0 = g.rd[0], rd1 = g.rd[1], rd2 = g.rd[2], dt = g.dt;
  const float by = ysl >= 0 ? pm.bE[1][j] : 0.f, cy = ysl >= 0 ? pm.cE[1][j] : 0.f;
  // TF/SF fold: this lane's packed weights (zero away from the y and z faces)
  int tfw[V];
#pragma unroll
  for (int q = 0; q < V; q++) tfw[q] = 0;
  if constexpr (TF) {
    const int* hd = reinterpret_cast<const int*>(fold);
#pragma unroll
    for (int q = 0; q < V; q++) tfw[q] = fold_pack(1, j, k + q, hd);
  }

  // prefetch depth, as in the H sweep; the paired variant with the folded
  // plane wave keeps one set (two spill 30 VGPRs; one set measured 5% faster)
  constexpr int PF = (V == 2 && TF) ? 1 : ((GY || V == 2) ? 2 : 3);
  PreE<P, T, CL> S[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    S[u].hx = S[u].hxy = S[u].hy = S[u].hz = S[u].hzy = S[u].ex = S[u].ey = S[u].ez = pzero<P>();
    S[u].hxz = S[u].hyz = rzero<T>();
    S[u].cls = 0;
  }
  ring_direct(t.i0 - 1);
  ring_direct(t.i0);
  ring_direct(t.i0 + 1);
#pragma unroll
  for (int u = 0; u < PF; u++) fetch(S[u], min(t.i0 + u, t.i1 - 1));  // unconditional, as in the H sweep

  auto step = [&](int i, PreE<P, T, CL>& Q) {
    const int b = i & 1;
    const int s0 = ring5(i - 1), s1 = ring5(i), s2 = ring5(i + 1), s3 = ring5(i + 2);
    // publish plane i (Hx) and plane i+2 (Hy, Hz); E operands to the own slots
    float v[V];
    unpackv(Q.hx, ihx, v);
#pragma unroll
    for (int q = 0; q < V; q++) sHx[b][ty + 1][c0 + q + 1] = v[q];
    unpackv(Q.hy, ihy, v);
#pragma unroll
    for (int q = 0; q < V; q++) rHy[s3][ty][c0 + q + 1] = v[q];
    unpackv(Q.hz, ihz, v);
#pragma unroll
    for (int q = 0; q < V; q++) rHz[s3][ty + 1][c0 + q] = v[q];
    if (yw) {
      unpackv(Q.hxy, ihx, v);
#pragma unroll
      for (int q = 0; q < V; q++) sHx[b][yrow][c0 + q + 1] = v[q];
      unpackv(Q.hzy, ihz, v);
#pragma unroll
      for (int q = 0; q < V; q++) rHz[s3][yrow][c0 + q] = v[q];
    }
    if (zl) {
      sHx[b][ty + 1][zcol] = cvt(Q.hxz, ihx);
      rHy[s3][ty][zcol] = cvt(Q.hyz, ihy);
    }
    unpackv(Q.ex, iex, v);
#pragma unroll
    for (int q = 0; q < V; q++) sEo[0][ty][c0 + q] = v[q];
    unpackv(Q.ey, iey, v);
#pragma unroll
    for (int q = 0; q < V; q++) sEo[1][ty][c0 + q] = v[q];
    unpackv(Q.ez, iez, v);
#pragma unroll
    for (int q = 0; q < V; q++) sEo[2][ty][c0 + q] = v[q];
#pragma unroll
    for (int q = 0; q < V; q++) sCl[ty][c0 + q] = (uint8_t)(Q.cls >> (8 * q));
    if constexpr (XZ) {
      if (xtile && (g.per[0] || (i + 2 >= xz.i0 && i < xz.i1))) {  // Hx and own E of plane i, Hy / Hz of i+2
        const int ii = GX ? nbr(i, 0, g.nx, g.per[0]) : i;
        const int i2n = GX ? nbr(i + 2, 0, g.nx, g.per[0]) : i + 2;
#pragma unroll
        for (int q = 0; q < V; q++) {
          ov(sHx[b][ty + 1][c0 + q + 1], 3, ii, jo, k + q);
          ov(rHy[s3][ty][c0 + q + 1], 4, i2n, jo, k + q);
          ov(rHz[s3][ty + 1][c0 + q], 5, i2n, jo, k + q);
          ov(sEo[0][ty][c0 + q], 0, ii, jo, k + q);
          ov(sEo[1][ty][c0 + q], 1, ii, jo, k + q);
          ov(sEo[2][ty][c0 + q], 2, ii, jo, k + q);
          if (yw) {
            ov(sHx[b][yrow][c0 + q + 1], 3, ii, jy, k + q);
            ov(rHz[s3][yrow][c0 + q], 5, i2n, jy, k + q);
          }
        }
        if (zl) {
          ov(sHx[b][ty + 1][zcol], 3, ii, jo, kz);
          ov(rHy[s3][ty][zcol], 4, i2n, jo, kz);
        }
      }
    }
    // psi and design-box operands of plane i before the prefetch
    float pz0[V], pz1[V], py0[V], py1[V], dv[V];
    bool over[V];
    const unsigned zb = psi_base_z(g, pm, i, j), yb = psi_base_y(g, pm, ysl, i, k);
#pragma unroll
    for (int q = 0; q < V; q++) {
      pz0[q] = pz1[q] = py0[q] = py1[q] = dv[q] = 0.f;
      if (!SP && inreg && zs[q] >= 0) {
        pz0[q] = pm.psiE[2][0][zb + zs[q]];
        pz1[q] = pm.psiE[2][1][zb + zs[q]];
      }
      if (!SP && GY && inreg && ysl >= 0) {
        py0[q] = pm.psiE[1][0][yb + q];
        py1[q] = pm.psiE[1][1][yb + q];
      }
      over[q] = DES && inreg && design_col[q] && i >= mt.d_i0 && i < mt.d_i0 + mt.d_ni;
      if (DES && over[q])
        dv[q] = mt.d_inv_eps[((long long)(i - mt.d_i0) * mt.d_nj + (j - mt.d_j0)) * mt.d_nk + (k + q - mt.d_k0)];
    }
    const int xsl = (GX && pm.on) ? slab_index(i, g.nx, pm.plo[0], pm.phi[0]) : -1;
    // this plane's fold factors, loaded before the barrier (whose wait covers
    // them) rather than inside the band branch
    float fu[6];
#pragma unroll
    for (int t6 = 0; t6 < 6; t6++) fu[t6] = 0.f;
    if constexpr (TF) {
#pragma unroll
      for (int t6 = 0; t6 < 6; t6++) fu[t6] = fold[kFoldStride * (1 + i) + t6];
    }
    fetch(Q, min(i + PF, t.i1 - 1));
    lds_barrier();
    if (inreg) {
      float nE[3][V];
      bool plain = true;
#pragma unroll
      for (int q = 0; q < V; q++) {
        const int c = c0 + q;
        const float ex0 = sEo[0][ty][c], ey0 = sEo[1][ty][c], ez0 = sEo[2][ty][c];
        const int cls = sCl[ty][c];
        const float hxo = sHx[b][ty + 1][c + 1];
        float dyHz = rd1 * d24(rHz[s1][ty][c], rHz[s1][ty + 1][c], rHz[s1][ty + 2][c], rHz[s1][ty + 3][c]);
        float dyHx = rd1 * d24(sHx[b][ty][c + 1], hxo, sHx[b][ty + 2][c + 1], sHx[b][ty + 3][c + 1]);
        float dzHy = rd2 * d24(rHy[s1][ty][c], rHy[s1][ty][c + 1], rHy[s1][ty][c + 2], rHy[s1][ty][c + 3]);
        float dzHx = rd2 * d24(sHx[b][ty + 1][c], hxo, sHx[b][ty + 1][c + 2], sHx[b][ty + 1][c + 3]);
        float dxHz = rd0 * d24(rHz[s0][ty + 1][c], rHz[s1][ty + 1][c], rHz[s2][ty + 1][c], rHz[s3][ty + 1][c]);
        float dxHy = rd0 * d24(rHy[s0][ty][c + 1], rHy[s1][ty][c + 1], rHy[s2][ty][c + 1], rHy[s3][ty][c + 1]);
        if constexpr (TF) {
          if (tfw[q]) {
            const int w = tfw[q];
            dyHz += fold_f(w, 0) * fu[0];   // Ex, y faces
            dzHy += fold_f(w, 6) * fu[1];   // Ex, z faces
            dzHx += fold_f(w, 12) * fu[2];  // Ey, z faces
            dxHz += fold_b(w, 24) * fu[3];  // Ey, x faces
            dxHy += fold_b(w, 25) * fu[4];  // Ez, x faces
            dyHx += fold_f(w, 18) * fu[5];  // Ez, y faces
          }
        }
        if (!SP && zs[q] >= 0) {  // z axis: slot 0 = Ex (D_z Hy), slot 1 = Ey (D_z Hx)
          cpml_step(pm.psiE[2][0], zb + zs[q], pz0[q], bz[q], cz[q], dzHy);
          cpml_step(pm.psiE[2][1], zb + zs[q], pz1[q], bz[q], cz[q], dzHx);
        }
        if (!SP && GX && xsl >= 0) {  // x axis: slot 0 = Ey (D_x Hz), slot 1 = Ez (D_x Hy)
          const unsigned qx = psi_base_x(g, pm, xsl, j, k) + q;
          cpml_apply(pm.psiE[0][0], qx, pm.bE[0][i], pm.cE[0][i], dxHz);
          cpml_apply(pm.psiE[0][1], qx, pm.bE[0][i], pm.cE[0][i], dxHy);
        }
        if (!SP && GY && ysl >= 0) {  // y axis: slot 0 = Ez (D_y Hx), slot 1 = Ex (D_y Hz)
          cpml_step(pm.psiE[1][0], yb + q, py0[q], by, cy, dyHx);
          cpml_step(pm.psiE[1][1], yb + q, py1[q], by, cy, dyHz);
        }
        const float crx = dyHz - dzHy, cry = dzHx - dxHz, crz = dxHy - dyHx;
        const ClassEntry& ce = tbl[cls];
        float ie0 = ce.inv_eps[0], ie1 = ce.inv_eps[1], ie2 = ce.inv_eps[2];
        if (over[q]) ie0 = ie1 = ie2 = dv[q];
        if (MAT && ce.special && !over[q]) {  // finished by k_ade; its E word keeps the old value here
          const long long bq =
              (long long)(i - ade.i0) * ade.bsx + (long long)(j - ade.j0) * ade.bsy + (k + q - ade.k0);
          ade.curl[0][bq] = crx;
          ade.curl[1][bq] = cry;
          ade.curl[2][bq] = crz;
          nE[0][q] = ex0;
          nE[1][q] = ey0;
          nE[2][q] = ez0;
          plain = false;
        } else if constexpr (SP) {
          // split parts: Ex (y, z), Ey (z, x), Ez (x, y); dispersive cells never sit in the layer
          const bool sx = GX && xsl >= 0;
          const float k1x = sx ? pm.bE[0][i] : 1.f, k2x = sx ? pm.cE[0][i] : 1.f;
          const float k1y = (GY && ysl >= 0) ? by : 1.f, k2y = (GY && ysl >= 0) ? cy : 1.f;
          const float k1z = zs[q] >= 0 ? bz[q] : 1.f, k2z = zs[q] >= 0 ? cz[q] : 1.f;
          const long long qs = plane_base<GX>(g, i) + own + q;
          nE[0][q] = split_step(pm.p2E[0], qs, ex0, k1y, k2y, dt * ie0 * dyHz, k1z, k2z, -dt * ie0 * dzHy);
          nE[1][q] = split_step(pm.p2E[1], qs, ey0, k1z, k2z, dt * ie1 * dzHx, k1x, k2x, -dt * ie1 * dxHz);
          nE[2][q] = split_step(pm.p2E[2], qs, ez0, k1x, k2x, dt * ie2 * dxHy, k1y, k2y, -dt * ie2 * dyHx);
        } else {
          nE[0][q] = ex0 + dt * ie0 * crx;
          nE[1][q] = ey0 + dt * ie1 * cry;
          nE[2][q] = ez0 + dt * ie2 * crz;
        }
      }
      const long long qq = plane_base<GX>(g, i) + own;
      if constexpr (XZ) {
        if (xtile) {  // new E of zone cells to the FP32 copy (a dispersive cell keeps its old value for k_ade)
#pragma unroll
          for (int q = 0; q < V; q++) {
            const int xq = exact_idx(xz, i, j, k + q);
            if (xq >= 0) {
              xz.f[0][xq] = nE[0][q];
              xz.f[1][xq] = nE[1][q];
              xz.f[2][xq] = nE[2][q];
            }
          }
        }
      }
      if constexpr (V == 2) {
        // a pair with a dispersive cell stores only its plain cells: k_ade
        // writes the other one (from the old value it reads)
        if (plain) {
          stv2(Ex, qq, nE[0], sc.se[0]);
          stv2(Ey, qq, nE[1], sc.se[1]);
          stv2(Ez, qq, nE[2], sc.se[2]);
        } else {
#pragma unroll
          for (int q = 0; q < V; q++) {
            if (tbl[sCl[ty][c0 + q]].special && !over[q]) continue;
            stv(Ex + q, qq, &nE[0][q], sc.se[0]);
            stv(Ey + q, qq, &nE[1][q], sc.se[1]);
            stv(Ez + q, qq, &nE[2][q], sc.se[2]);
          }
        }
      } else {
        if (plain) {
          stv(Ex, qq, nE[0], sc.se[0]);
          stv(Ey, qq, nE[1], sc.se[1]);
          stv(Ez, qq, nE[2], sc.se[2]);
        }
      }
    }
  };
  int i = t.i0;
  for (; i + PF - 1 < t.i1; i += PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) step(i + u, S[u]);
  }
#pragma unroll
  for (int u = 0; u < PF - 1; u++)
    if (i + u < t.i1) step(i + u, S[u]);
}

// ======================================================== small kernels
template <int STORE>
__global__ void k_soft(Fields<STORE> f, Scales sc, Soft so) {
  const int t = threadIdx.x;
  if (t >= so.n) return;
  const int c = so.comp[t];
  const long long q = so.q[t];
  const int xq = STORE == 1 ? exact_idx(sc.xz, so.ijk[t][0], so.ijk[t][1], so.ijk[t][2]) : -1;
  stx(f.E[c], q, ldx(f.E[c], q, sc.ie[c], sc.xz, c, xq) - so.val[t], sc.se[c], sc.xz, c, xq);
}

template <int STORE>
__device__ float fval(const Grid& g, const Fields<STORE>& f, const Scales& sc, int comp, int i, int j, int k) {
  const int ii = nbr(i, 0, g.nx, g.per[0]), jj = nbr(j, 0, g.ny, g.per[1]), kk = nbr(k, 0, g.nz, g.per[2]);
  if (ii < 0 || jj < 0 || kk < 0) return 0.f;
  const long long q = ii * g.sx + jj * g.sy + kk;
  const int xq = STORE == 1 ? exact_idx(sc.xz, ii, jj, kk) : -1;
  return comp < 3 ? ldx(f.E[comp], q, sc.ie[comp], sc.xz, comp, xq)
                  : ldx(f.H[comp - 3], q, sc.ih[comp - 3], sc.xz, comp, xq);
}

// Reference sensor interpolation (sensor.h:64-71) on the cell lattice.
template <int STORE>
__device__ float sensor_val(const Grid& g, const Fields<STORE>& f, const Scales& sc, int comp, int x, int y,
                            int z) {
  const int zM = (z == 0) ? z + 1 : z - 1;
  switch (comp) {
    case 0:
      return 0.25f * (fval(g, f, sc, 0, x, y - 1, zM) + fval(g, f, sc, 0, x, y - 1, z) +
                      fval(g, f, sc, 0, x, y, zM) + fval(g, f, sc, 0, x, y, z));
    case 1:
      return 0.25f * (fval(g, f, sc, 1, x - 1, y, zM) + fval(g, f, sc, 1, x - 1, y, z) +
                      fval(g, f, sc, 1, x, y, zM) + fval(g, f, sc, 1, x, y, z));
    case 2:
      return 0.25f * (fval(g, f, sc, 2, x - 1, y - 1, z) + fval(g, f, sc, 2, x - 1, y, z) +
                      fval(g, f, sc, 2, x, y - 1, z) + fval(g, f, sc, 2, x, y, z));
    case 3: return 0.5f * (fval(g, f, sc, 3, x - 1, y, z) + fval(g, f, sc, 3, x, y, z));
    case 4: return 0.5f * (fval(g, f, sc, 4, x, y - 1, z) + fval(g, f, sc, 4, x, y, z));
    default: return 0.5f * (fval(g, f, sc, 5, x, y, zM) + fval(g, f, sc, 5, x, y, z));
  }
}

template <int STORE>
// One thread per (probe, DFT bin): every thread of a probe reads its value
// (a few cached loads), the bin-0 thread stores it to the series, and each
// thread advances its own FP64 accumulator, so the sincos of the bins run in
// parallel rather than one after another (64 bins: 26 -> 5 us per step).
__global__ void k_record(Grid g, Fields<STORE> f, Scales sc, Probes pr) {
  const int nb = pr.nfreq > 0 ? pr.nfreq : 1;
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (tid >= (long long)pr.n * nb) return;
  const int p = (int)(tid / nb), fq = (int)(tid % nb);
  const ProbeDesc d = pr.desc[p];
  const float v = d.kind == 0 ? fval(g, f, sc, d.comp, d.i, d.j, d.k) : sensor_val(g, f, sc, d.comp, d.i, d.j, d.k);
  if (fq == 0) pr.series[(long long)pr.row * pr.n + p] = v;
  if (pr.nfreq > 0) {
    const double t = d.comp < 3 ? (double)(pr.step + 1) * pr.dt : ((double)pr.step + 0.5) * pr.dt;
    double* acc = pr.dft + ((long long)p * pr.nfreq + fq) * 2;
    double s, c;
    sincos(pr.omega[fq] * t, &s, &c);
    acc[0] += (double)v * c * pr.dt;
    acc[1] += (double)v * s * pr.dt;
  }
}

__global__ void k_to_float_h(const __half* src, float inv, float* dst, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = __half2float(src[q]) * inv;
}
__global__ void k_from_float_h(const float* src, float s, __half* dst, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    dst[q] = __float2half_rn(src[q] * s);
}

// ============================================================ launchers
// First failed or refused launch of this thread since the last
// take_launch_error (kernel name and reason).
static thread_local char t_launch_err[256];
static thread_local bool t_launch_failed = false;

void fail_launch(const char* kernel, const char* why) {
  if (t_launch_failed) return;  // keep the first
  snprintf(t_launch_err, sizeof(t_launch_err), "launch of %s failed: %s", kernel, why);
  t_launch_failed = true;
}

void note_launch(const char* kernel) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail_launch(kernel, hipGetErrorString(e));
}

int take_launch_error(char* msg, int cap) {
  if (!t_launch_failed) return 0;
  if (msg && cap > 0) snprintf(msg, (size_t)cap, "%s", t_launch_err);
  t_launch_failed = false;
  return 1;
}

static void check_launch(const char* what) { note_launch(what); }

// Blocks of one launch that fit on the device at once.
template <typename K>
static int resident_blocks(K kernel, int threads) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 512;
    cus = p.multiProcessorCount;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  return per * cus;
}

// Segment length along x: a block marches seg planes after a warm-up of
// about kWarm planes (three ring planes loaded synchronously, then the
// two-plane prefetch filling), and the launch runs in ceil(blocks / slots)
// rounds.  Minimise rounds * (seg + kWarm); *cost receives the minimum.
static int march_seg(const Region* regs, int nreg, int slots, int bz, int by, long long* cost_out = nullptr) {
  constexpr int kWarm = 6;
  int maxlen = 1;
  for (int t = 0; t < nreg; t++) {
    const Region& r = regs[t];
    if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
    maxlen = std::max(maxlen, r.i1 - r.i0);
  }
  int best = maxlen;
  long long best_cost = -1;
  for (int nseg = 1; nseg <= maxlen; nseg++) {
    const int seg = (maxlen + nseg - 1) / nseg;
    if (nseg > 1 && seg == (maxlen + nseg - 2) / (nseg - 1)) continue;  // same seg as nseg-1
    long long blocks = 0;
    for (int t = 0; t < nreg; t++) {
      const Region& r = regs[t];
      if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
      blocks += (long long)((r.k1 - r.k0 + bz - 1) / bz) * ((r.j1 - r.j0 + by - 1) / by) *
                ((r.i1 - r.i0 + seg - 1) / seg);
    }
    const long long rounds = (blocks + slots - 1) / slots;
    const long long cost = rounds * (seg + kWarm);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = seg;
    }
    if (seg <= 4) break;
  }
  if (cost_out) *cost_out = best_cost;
  return best;
}

// Tile height of a one-cell-per-lane sweep: 16 rows (one 1024-lane block per
// CU) unless the 8-row launch (two blocks per CU) has a lower march cost.  A
// block of either height keeps its CU equally busy per plane, so the costs
// compare directly.
static bool use_16_rows(int slots8, int slots16, const Region* regs, int nreg, int bz, int& seg8, int& seg16) {
  long long c8 = 0, c16 = 0;
  seg8 = march_seg(regs, nreg, slots8, bz, 8, &c8);
  seg16 = march_seg(regs, nreg, slots16, bz, 16, &c16);
  return c16 <= c8;
}

template <int STORE, int V, bool SP, bool XZ, bool TF, bool W = false>
static void upd_h(const Grid& g, const void* fields, const Scales& sc, const Cpml& pm, const Region* regs, int nreg,
                  const float* fold, hipStream_t st) {
  const Fields<STORE>& f = *static_cast<const Fields<STORE>*>(fields);
  int seg8 = 0, seg16 = 0;
  if constexpr (V == 1) {
    static const int slots8 = resident_blocks(k_update_h<STORE, 3, V, SP, 8, XZ, TF, W>, BZ * 8);
    static const int slots16 = resident_blocks(k_update_h<STORE, 3, V, SP, 16, XZ, TF, W>, BZ * 16);
    if (use_16_rows(slots8, slots16, regs, nreg, BZ, seg8, seg16)) {
      MarchSet rg = make_march(regs, nreg, seg16, BZ, 16);
      if (rg.total <= 0) return;
      hipLaunchKernelGGL((k_update_h<STORE, 3, V, SP, 16, XZ, TF, W>), dim3(rg.per * NXCD), dim3(BZ, 16, 1), 0, st, g, f,
                         sc, pm, rg, fold);
      check_launch("k_update_h");
      return;
    }
  } else {
    static const int slots = resident_blocks(k_update_h<STORE, 3, V, SP, 8, XZ, TF, W>, BZ * 8);
    seg8 = march_seg(regs, nreg, slots, BZ * V, 8);
  }
  MarchSet rg = make_march(regs, nreg, seg8, BZ * V, 8);
  if (rg.total <= 0) return;
  hipLaunchKernelGGL((k_update_h<STORE, 3, V, SP, 8, XZ, TF, W>), dim3(rg.per * NXCD), dim3(BZ, 8, 1), 0, st, g, f, sc, pm,
                     rg, fold);
  check_launch("k_update_h");
}

template <int STORE, int V, bool SP, bool XZ, bool TF, bool MAT = true, bool W = false>
static void upd_e(const Grid& g, const void* fields, const Scales& sc, const Cpml& pm, const Materials& mt,
                  const Ade& ade, const Region* regs, int nreg, const float* fold, hipStream_t st) {
  const Fields<STORE>& f = *static_cast<const Fields<STORE>*>(fields);
  int seg8 = 0, seg16 = 0;
  if constexpr (V == 1) {
    static const int slots8 = resident_blocks(k_update_e<STORE, 3, V, SP, 8, XZ, TF, MAT, W>, BZ * 8);
    static const int slots16 = resident_blocks(k_update_e<STORE, 3, V, SP, 16, XZ, TF, MAT, W>, BZ * 16);
    if (use_16_rows(slots8, slots16, regs, nreg, BZ, seg8, seg16)) {
      MarchSet rg = make_march(regs, nreg, seg16, BZ, 16);
      if (rg.total <= 0) return;
      hipLaunchKernelGGL((k_update_e<STORE, 3, V, SP, 16, XZ, TF, MAT, W>), dim3(rg.per * NXCD), dim3(BZ, 16, 1), 0, st, g,
                         f, sc, pm, mt, ade, rg, fold);
      check_launch("k_update_e");
      return;
    }
  } else {
    static const int slots = resident_blocks(k_update_e<STORE, 3, V, SP, 8, XZ, TF, MAT, W>, BZ * 8);
    seg8 = march_seg(regs, nreg, slots, BZ * V, 8);
  }
  MarchSet rg = make_march(regs, nreg, seg8, BZ * V, 8);
  if (rg.total <= 0) return;
  hipLaunchKernelGGL((k_update_e<STORE, 3, V, SP, 8, XZ, TF, MAT, W>), dim3(rg.per * NXCD), dim3(BZ, 8, 1), 0, st, g, f, sc,
                     pm, mt, ade, rg, fold);
  check_launch("k_update_e");
}

// FP16 storage on an even nz runs two z cells per lane (__half2 words);
// a design box (FP32 inverse design) keeps one cell per lane.
static bool pair_cells(int store, const Grid& g, const Materials* mt) {
  return store == 1 && g.nz % 2 == 0 && !(mt && mt->d_on);
}
// pairs start at an even k: every region must too (the __half2 words)
static bool even_k(const Region* regs, int nreg) {
  for (int r = 0; r < nreg; r++)
    if (regs[r].k0 & 1) return false;
  return true;
}

// The sweep variant of a configuration: storage, cells per lane, layer
// kind, FP32 zone (its own variants, so that the other variants carry none of
// its registers) and a folded plane wave along x (never with the zone).
template <bool TF>
static void update_h_go(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                        const Region* regs, int nreg, const float* fold, hipStream_t st) {
  const bool xz = store == 1 && sc.xz.on;
  const bool pairs = pair_cells(store, g, nullptr) && even_k(regs, nreg);
  if (wide_grid(g)) {  // 64-bit cell indices (no split layer or zone there: dtm_create refuses them)
    if (store == 0) upd_h<0, 1, false, false, TF, true>(g, fields, sc, pm, regs, nreg, fold, st);
    else if (pairs) upd_h<1, 2, false, false, TF, true>(g, fields, sc, pm, regs, nreg, fold, st);
    else upd_h<1, 1, false, false, TF, true>(g, fields, sc, pm, regs, nreg, fold, st);
    return;
  }
  if (pm.split) {  // split-field layer: one cell per lane
    if (store == 0) upd_h<0, 1, true, false, TF>(g, fields, sc, pm, regs, nreg, fold, st);
    else if (xz) upd_h<1, 1, true, true, false>(g, fields, sc, pm, regs, nreg, fold, st);
    else upd_h<1, 1, true, false, TF>(g, fields, sc, pm, regs, nreg, fold, st);
  } else if (store == 0) {
    upd_h<0, 1, false, false, TF>(g, fields, sc, pm, regs, nreg, fold, st);
  } else if (xz) {
    if (pairs) upd_h<1, 2, false, true, false>(g, fields, sc, pm, regs, nreg, fold, st);
    else upd_h<1, 1, false, true, false>(g, fields, sc, pm, regs, nreg, fold, st);
  } else if (pairs) {
    upd_h<1, 2, false, false, TF>(g, fields, sc, pm, regs, nreg, fold, st);
  } else {
    upd_h<1, 1, false, false, TF>(g, fields, sc, pm, regs, nreg, fold, st);
  }
}

template <bool TF>
static void update_e_go(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                        const Materials& mt, const Ade& ade, const Region* regs, int nreg, const float* fold,
                        hipStream_t st) {
  const bool xz = store == 1 && sc.xz.on;
  const bool pairs = pair_cells(store, g, &mt) && even_k(regs, nreg);
  const bool mat = ade.on || mt.d_on;  // else the variants without the dispersive / design branches
  if (wide_grid(g)) {  // 64-bit cell indices, the material branches kept
    if (store == 0) upd_e<0, 1, false, false, TF, true, true>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else if (pairs) upd_e<1, 2, false, false, TF, true, true>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<1, 1, false, false, TF, true, true>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    return;
  }
  if (pm.split) {
    if (store == 0) upd_e<0, 1, true, false, TF>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else if (xz) upd_e<1, 1, true, true, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<1, 1, true, false, TF>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  } else if (store == 0) {
    if (mat) upd_e<0, 1, false, false, TF>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<0, 1, false, false, TF, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  } else if (xz) {
    if (pairs) upd_e<1, 2, false, true, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<1, 1, false, true, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  } else if (pairs) {
    if (mat) upd_e<1, 2, false, false, TF>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<1, 2, false, false, TF, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  } else {
    if (mat) upd_e<1, 1, false, false, TF>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
    else upd_e<1, 1, false, false, TF, false>(g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  }
}

// With an FP32 zone only a column of cells (the zone's y and z range plus
// the 3 cells its stencils reach, widened to the tile grid, every x plane)
// runs the zone variants; the rest of the grid runs the plain ones as up to
// four boxes around it (y slabs, then z slabs inside the column's y range),
// all of full x extent, so every launch keeps the long x segments of the
// whole grid.  Both launches of a sweep write disjoint cells and read only
// the other family of fields, so their order does not matter and they may
// run concurrently.  A periodic axis keeps the column at its full extent on
// that axis.
static bool zone_split(const Grid& g, const Exact& x, const Region* regs, int nreg, Region* outer, int& nouter,
                       Region& mid) {
  if (nreg != 1) return false;
  const Region& R = regs[0];
  int j0 = g.per[1] ? R.j0 : max(R.j0, x.j0 - 3), j1 = g.per[1] ? R.j1 : min(R.j1, x.j1 + 3);
  int k0 = g.per[2] ? R.k0 : max(R.k0, x.k0 - 3), k1 = g.per[2] ? R.k1 : min(R.k1, x.k1 + 3);
  if (j0 >= j1 || k0 >= k1) return false;
  // the column on the region's tile grid (rows in 16s, z in 128s from its
  // origin: every tile shape of the plain launch), so the outer boxes hold
  // whole tiles of the one-box tiling and add no partial tiles; z bounds are
  // then even, as the paired FP16 sweeps need (two z cells per lane from an
  // even k)
  j0 = R.j0 + (j0 - R.j0) / 16 * 16;
  j1 = min(R.j1, R.j0 + (j1 - R.j0 + 15) / 16 * 16);
  k0 = R.k0 + (k0 - R.k0) / 128 * 128;
  k1 = min(R.k1, R.k0 + (k1 - R.k0 + 127) / 128 * 128);
  k0 &= ~1;
  k1 = min(R.k1, (k1 + 1) & ~1);
  nouter = 0;
  auto add = [&](int a0, int a1, int b0, int b1) {
    if (a0 < a1 && b0 < b1) outer[nouter++] = Region{R.i0, R.i1, a0, a1, b0, b1};
  };
  add(R.j0, j0, R.k0, R.k1);
  add(j1, R.j1, R.k0, R.k1);
  add(j0, j1, R.k0, k0);
  add(j0, j1, k1, R.k1);
  mid = Region{R.i0, R.i1, j0, j1, k0, k1};
  return true;
}

// The zone's column and the outer boxes write disjoint cells and read only
// the other family of fields: with a side stream they run concurrently.
static void side_fork(const SideStream* side, hipStream_t st) {
  if (!side) return;
  hipError_t e = hipEventRecord(side->fork, st);
  if (e == hipSuccess) e = hipStreamWaitEvent(side->st, side->fork, 0);
  if (e != hipSuccess) fail_launch("side stream fork", hipGetErrorString(e));
}
static void side_join(const SideStream* side, hipStream_t st) {
  if (!side) return;
  hipError_t e = hipEventRecord(side->join, side->st);
  if (e == hipSuccess) e = hipStreamWaitEvent(st, side->join, 0);
  if (e != hipSuccess) fail_launch("side stream join", hipGetErrorString(e));
}

void launch_update_h(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                     const Region* regs, int nreg, hipStream_t st, const float* fold, const SideStream* side) {
  Region outer[4], mid;
  int nouter = 0;
  if (store == 1 && sc.xz.on && zone_split(g, sc.xz, regs, nreg, outer, nouter, mid)) {
    Scales plain = sc;
    plain.xz.on = 0;
    side_fork(side, st);
    update_h_go<false>(store, g, fields, sc, pm, &mid, 1, nullptr, side ? side->st : st);
    if (nouter) update_h_go<false>(store, g, fields, plain, pm, outer, nouter, nullptr, st);
    side_join(side, st);
    return;
  }
  if (fold && !(store == 1 && sc.xz.on))
    update_h_go<true>(store, g, fields, sc, pm, regs, nreg, fold, st);
  else
    update_h_go<false>(store, g, fields, sc, pm, regs, nreg, nullptr, st);
}

void launch_update_e(int store, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                     const Materials& mt, const Ade& ade, const Region* regs, int nreg, hipStream_t st,
                     const float* fold, const SideStream* side) {
  Region outer[4], mid;
  int nouter = 0;
  if (store == 1 && sc.xz.on && zone_split(g, sc.xz, regs, nreg, outer, nouter, mid)) {
    Scales plain = sc;
    plain.xz.on = 0;
    side_fork(side, st);
    update_e_go<false>(store, g, fields, sc, pm, mt, ade, &mid, 1, nullptr, side ? side->st : st);
    if (nouter) update_e_go<false>(store, g, fields, plain, pm, mt, ade, outer, nouter, nullptr, st);
    side_join(side, st);
    return;
  }
  if (fold && !(store == 1 && sc.xz.on))
    update_e_go<true>(store, g, fields, sc, pm, mt, ade, regs, nreg, fold, st);
  else
    update_e_go<false>(store, g, fields, sc, pm, mt, ade, regs, nreg, nullptr, st);
}

// One launch over the union of the face bands (k_tfsf gives each cell one
// writer); the host fills each band's first thread index and its overlap
// mask.
template <int STORE, int MODE>
static void tfsf_go(const Grid& g, const void* fields, const Scales& sc, const Cpml& pm, const Materials& mt,
                    const Ade& ade, const Tfsf& tf, const TfsfBands& bands, double tinc, hipStream_t st) {
  const Fields<STORE>& f = *static_cast<const Fields<STORE>*>(fields);
  TfsfBands b = bands;
  long long n = 0;
  b.start[0] = 0;
  for (int r = 0; r < b.n; r++) {
    const Region& R = b.r[r];
    n += ((long long)(R.i1 - R.i0) * (R.j1 - R.j0) * (R.k1 - R.k0) + 63) / 64 * 64;  // wavefront-aligned
    b.start[r + 1] = (unsigned)std::min<long long>(n, 0xffffffffLL);
    b.ovl[r] = 0;
    for (int q = 0; q < b.n; q++) {
      const Region& Q = b.r[q];
      if (q != r && R.i0 < Q.i1 && Q.i0 < R.i1 && R.j0 < Q.j1 && Q.j0 < R.j1 && R.k0 < Q.k1 && Q.k0 < R.k1)
        b.ovl[r] |= 1u << q;
    }
  }
  if (n <= 0) return;
  if (n >= (1LL << 31)) {  // dtm_create rejects such boxes; never index past 2^31 threads
    fail_launch("k_tfsf", "TF/SF face bands hold 2^31 cells or more");
    return;
  }
  const dim3 grid((unsigned)((n + 255) / 256));
  if (tf.tab_axis >= 0)
    hipLaunchKernelGGL((k_tfsf<STORE, MODE, 1>), grid, dim3(256), 0, st, g, f, sc, pm, mt, ade, tf, b, tinc);
  else if (tf.ptab)
    hipLaunchKernelGGL((k_tfsf<STORE, MODE, 2>), grid, dim3(256), 0, st, g, f, sc, pm, mt, ade, tf, b, tinc);
  else
    hipLaunchKernelGGL((k_tfsf<STORE, MODE, 0>), grid, dim3(256), 0, st, g, f, sc, pm, mt, ade, tf, b, tinc);
  check_launch("k_tfsf");
}

void launch_tfsf_tables(const Tfsf& tf, double* out, long long n0, int nsteps, double dt, hipStream_t st) {
  const long long n = (long long)nsteps * 2 * tf.tab_n;
  hipLaunchKernelGGL(k_tfsf_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tf, out, n0, nsteps, dt);
  check_launch("k_tfsf_table");
}

// Per-plane factors of the folded TF/SF source (plane wave along x) for the
// half-steps of a table batch: out[(s*2 + half)*(nx+1)*kFoldStride] holds the
// header, then plane i's six factors at +kFoldStride*(1+i).  tables: the
// batch's waveform tables (k_tfsf_table).  With T(h) = P-less table value at
// x half-coordinate h, xin(h) = h inside the box on x and
// X(h) = sum over the neighbours h+o across an x face of wt_o (+-1) T(h+o):
//   H half: rd1 xin(2i+1) P[Ez] T(2i+1), rd2 xin(2i+1) P[Ey] T(2i+1),
//           rd2 xin(2i) P[Ex] T(2i),     rd0 P[Ez] X(2i), rd0 P[Ey] X(2i),
//           rd1 xin(2i) P[Ex] T(2i);
//   E half: rd1 xin(2i) P[Hz] T(2i),     rd2 xin(2i) P[Hy] T(2i),
//           rd2 xin(2i+1) P[Hx] T(2i+1), rd0 P[Hz] X(2i+1), rd0 P[Hy] X(2i+1),
//           rd1 xin(2i+1) P[Hx] T(2i+1).
__global__ __launch_bounds__(256) void k_tfsf_fold(Tfsf tf, const double* tables, float* out, int nsteps, int nx,
                                                   float rd0, float rd1, float rd2) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = nx + 1;
  if (idx >= nsteps * 2 * per) return;
  const int sh = idx / per, p = idx % per;  // sh = step * 2 + half
  const int half = sh & 1;
  float* o = out + (long long)sh * per * kFoldStride;
  if (p == 0) {  // header
    int* hd = reinterpret_cast<int*>(o);
    hd[0] = tf.hlo[1];
    hd[1] = tf.hhi[1];
    hd[2] = tf.hlo[2];
    hd[3] = tf.hhi[2];
    hd[4] = tf.sf_inside;
    hd[5] = hd[6] = hd[7] = 0;
    return;
  }
  const int i = p - 1;
  const double* T = tables + (long long)sh * tf.tab_n;
  auto Tv = [&](int h) { return T[min(max(h - tf.tab_h0, 0), tf.tab_n - 1)]; };
  auto xin = [&](int h) { return h >= tf.hlo[0] && h <= tf.hhi[0] ? 1.0 : 0.0; };
  auto X = [&](int h) {
    const bool pin = h >= tf.hlo[0] && h <= tf.hhi[0];
    const int ho[4] = {-3, -1, 1, 3};
    const double wt[4] = {1.0, -27.0, 27.0, -1.0};
    double acc = 0.0;
    for (int q = 0; q < 4; q++) {
      const int hq = h + ho[q];
      const bool qin = hq >= tf.hlo[0] && hq <= tf.hhi[0];
      if (qin == pin) continue;
      const double v = Tv(hq);
      acc += wt[q] * ((tf.sf_inside ? qin : !qin) ? v : -v);
    }
    return acc;
  };
  const int e = 2 * i, od = 2 * i + 1;
  float* u = o + kFoldStride * p;
  if (half == 0) {  // neighbours are E components (0, 1, 2)
    u[0] = (float)(rd1 * xin(od) * tf.P[2] * Tv(od));
    u[1] = (float)(rd2 * xin(od) * tf.P[1] * Tv(od));
    u[2] = (float)(rd2 * xin(e) * tf.P[0] * Tv(e));
    u[3] = (float)(rd0 * tf.P[2] * X(e));
    u[4] = (float)(rd0 * tf.P[1] * X(e));
    u[5] = (float)(rd1 * xin(e) * tf.P[0] * Tv(e));
  } else {          // neighbours are H components (3, 4, 5)
    u[0] = (float)(rd1 * xin(e) * tf.P[5] * Tv(e));
    u[1] = (float)(rd2 * xin(e) * tf.P[4] * Tv(e));
    u[2] = (float)(rd2 * xin(od) * tf.P[3] * Tv(od));
    u[3] = (float)(rd0 * tf.P[5] * X(od));
    u[4] = (float)(rd0 * tf.P[4] * X(od));
    u[5] = (float)(rd1 * xin(od) * tf.P[3] * Tv(od));
  }
  u[6] = u[7] = 0.f;
}

void launch_tfsf_fold(const Tfsf& tf, const double* tables, float* out, int nsteps, const Grid& g,
                      hipStream_t st) {
  const int n = nsteps * 2 * (g.nx + 1);
  hipLaunchKernelGGL(k_tfsf_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tf, tables, out, nsteps,
                     g.nx, g.rd[0], g.rd[1], g.rd[2]);
  check_launch("k_tfsf_fold");
}

// Incident values of every component on the point table's box at time t:
// one thread per value, so the FP64 transcendentals of a dipole run across
// the chip instead of in the band threads one after another.
__global__ __launch_bounds__(256) void k_tfsf_points(Tfsf tf, double t) {
  const long long n = 6LL * tf.ptab_n[0] * tf.ptab_n[1] * tf.ptab_n[2];
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int p2 = (int)(q % tf.ptab_n[2]);
  const long long r = q / tf.ptab_n[2];
  const int p1 = (int)(r % tf.ptab_n[1]);
  const int p0 = (int)((r / tf.ptab_n[1]) % tf.ptab_n[0]);
  const int c = (int)(r / tf.ptab_n[1] / tf.ptab_n[0]);
  tf.ptab[q] = incident(tf, c, tf.ptab_h0[0] + p0, tf.ptab_h0[1] + p1, tf.ptab_h0[2] + p2, t);
}

void launch_tfsf(int store, int half, const Grid& g, const void* fields, const Scales& sc, const Cpml& pm,
                 const Materials& mt, const Ade& ade, const Tfsf& tf, const TfsfBands& bands, double t_inc,
                 hipStream_t st) {
  if (tf.ptab && tf.tab_axis < 0) {
    const long long n = 6LL * tf.ptab_n[0] * tf.ptab_n[1] * tf.ptab_n[2];
    hipLaunchKernelGGL(k_tfsf_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tf, t_inc);
    check_launch("k_tfsf_points");
  }
  if (store == 0 && half == 0) tfsf_go<0, 0>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
  if (store == 0 && half == 1) tfsf_go<0, 1>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
  if (store == 0 && half == 2) tfsf_go<0, 2>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
  if (store == 1 && half == 0) tfsf_go<1, 0>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
  if (store == 1 && half == 1) tfsf_go<1, 1>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
  if (store == 1 && half == 2) tfsf_go<1, 2>(g, fields, sc, pm, mt, ade, tf, bands, t_inc, st);
}

void launch_ade(int store, const Grid& g, const void* fsrc, const void* fdst, const Scales& sc, const Materials& mt,
                const Ade& ade, hipStream_t st) {
  if (!ade.on) return;
  const long long nb = (long long)ade.ni * ade.nj * ade.nk;
  const unsigned blocks = (unsigned)std::min<long long>((nb + 255) / 256, 65536);
  if (store == 0)
    hipLaunchKernelGGL(k_ade<0>, dim3(blocks), dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fsrc),
                       *static_cast<const Fields<0>*>(fdst), sc, mt, ade);
  else
    hipLaunchKernelGGL(k_ade<1>, dim3(blocks), dim3(256), 0, st, g, *static_cast<const Fields<1>*>(fsrc),
                       *static_cast<const Fields<1>*>(fdst), sc, mt, ade);
  check_launch("k_ade");
}

void launch_soft(int store, const Grid&, const void* fields, const Scales& sc, const Soft& so, hipStream_t st) {
  if (so.n <= 0) return;
  if (store == 0)
    hipLaunchKernelGGL(k_soft<0>, dim3(1), dim3(64), 0, st, *static_cast<const Fields<0>*>(fields), sc, so);
  else
    hipLaunchKernelGGL(k_soft<1>, dim3(1), dim3(64), 0, st, *static_cast<const Fields<1>*>(fields), sc, so);
  check_launch("k_soft");
}

void launch_record(int store, const Grid& g, const void* fields, const Scales& sc, const Probes& pr, hipStream_t st) {
  if (pr.n <= 0) return;
  const long long threads = (long long)pr.n * (pr.nfreq > 0 ? pr.nfreq : 1);
  const int nb = (int)((threads + 63) / 64);
  if (store == 0)
    hipLaunchKernelGGL(k_record<0>, dim3(nb), dim3(64), 0, st, g, *static_cast<const Fields<0>*>(fields), sc, pr);
  else
    hipLaunchKernelGGL(k_record<1>, dim3(nb), dim3(64), 0, st, g, *static_cast<const Fields<1>*>(fields), sc, pr);
  check_launch("k_record");
}

void launch_to_float(int store, const void* src, float inv, float* dst, long long n, hipStream_t st) {
  if (store == 0) {
    if (hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
      fail_launch("field copy", hipGetErrorString(hipGetLastError()));
    return;
  }
  hipLaunchKernelGGL(k_to_float_h, dim3(1024), dim3(256), 0, st, static_cast<const __half*>(src), inv, dst, n);
  check_launch("k_to_float_h");
}

void launch_from_float(int store, const float* src, float s, void* dst, long long n, hipStream_t st) {
  if (store == 0) {
    if (hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
      fail_launch("field copy", hipGetErrorString(hipGetLastError()));
    return;
  }
  hipLaunchKernelGGL(k_from_float_h, dim3(1024), dim3(256), 0, st, src, s, static_cast<__half*>(dst), n);
  check_launch("k_from_float_h");
}

const char* kernel_variant_name(int store) {
  return store == 0 ? "k_update_{h,e}<fp32>" : "k_update_{h,e}<fp16>";
}

}  // namespace dtm
