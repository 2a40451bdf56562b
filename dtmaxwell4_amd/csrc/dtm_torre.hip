// dtm_torre.hip — fused single-pass time step ("torre" kernel) for gfx950.
//
// One launch advances H^{n-1/2} -> H^{n+1/2} and E^n -> E^{n+1} for a set of
// regions, reading the old fields once and writing the new ones once
// (48 B per cell-update in FP32 instead of the 72 B of two separate sweeps).
// This is the temporal fusion of the reference's DiamondTorre, where a torre
// advances its body diamond (H) and then its nose diamond (E) while sliding
// one rag along x per step (genAlg.py:385-499, diamond.cu:72-262); here a
// block slides along x and the E update trails the H update by two planes,
// the reach of the (2,4) stencil.
//
// Block = 64 lanes (z) x TR=16 rows (y).  Thread (ty, tx) computes H at
// (j0-1+ty, k0-1+tx) on every plane and, for the owned interior
// ty in [1,14), tx in [1,62), E at (j0-1+ty, k0-1+tx) two planes behind.
// Halo H values are recomputed by neighbouring blocks from the same old
// fields, so fields and CPML psi are double-buffered (read old, write new)
// and blocks need no synchronisation with each other.
//
// LDS: the E tile of the current plane (19 x 67 per component, 2 slots) for
// the y/z stencil of H, and a 4-slot ring of the 16 x 64 H planes for the
// y/z stencil of E.  One LDS-only barrier per plane; the next plane's global
// loads are issued before it.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdio.h>

#include <algorithm>

#include "dtm_internal.h"

namespace dtm {

namespace {

constexpr int TZ = 64;            // lanes per row (one wavefront)
constexpr int TR = 16;            // thread rows
constexpr int OZ = TZ - 3;        // owned z per tile (61)
constexpr int OR = TR - 3;        // owned rows per tile (13)
constexpr int ER = TR + 3;        // E tile rows (19): j0-3 .. j0+15
constexpr int EC = TZ + 3;        // E tile cols (67): k0-3 .. k0+63
constexpr int ETILE = ER * EC;    // 1273
constexpr int NXCD = 8;

__device__ __forceinline__ float ld(const float* F, long long q, float) { return F[q]; }
__device__ __forceinline__ float ld(const __half* F, long long q, float inv) { return __half2float(F[q]) * inv; }
__device__ __forceinline__ void st(float* F, long long q, float v, float) { F[q] = v; }
__device__ __forceinline__ void st(__half* F, long long q, float v, float s) { F[q] = __float2half_rn(v * s); }

__device__ __forceinline__ int nbr(int v, int n, int per) {
  if (v >= 0 && v < n) return v;
  if (!per) return -1;
  return v < 0 ? v + n : v - n;
}
__device__ __forceinline__ float d24(float f0, float f1, float f2, float f3) { return 27.0f * (f2 - f1) - (f3 - f0); }
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ int slab_index(int v, int n, int plo, int phi) {
  if (v < plo) return v;
  if (v >= n - phi) return plo + v - (n - phi);
  return -1;
}
__device__ __forceinline__ long long psi_index(const Grid& g, const Cpml& p, int a, int s, int i, int j, int k) {
  const int P = p.plo[a] + p.phi[a];
  if (a == 0) return ((long long)s * g.ny + j) * g.nz + k;
  if (a == 1) return ((long long)i * P + s) * g.nz + k;
  return ((long long)i * g.ny + j) * P + s;
}

struct TorreTile {
  int r, i0, i1, j0, k0;
};

__device__ __forceinline__ bool torre_locate(const TorreSet& ts, TorreTile& t) {
  const int b = blockIdx.x;
  const int L = (b % NXCD) * ts.per + b / NXCD;  // XCD-contiguous logical id
  if (L >= ts.total) return false;
  int r = 0;
  while (r + 1 < ts.n && L >= ts.start[r + 1]) r++;
  int loc = L - ts.start[r];
  const int bk = loc % ts.nbk[r];
  loc /= ts.nbk[r];
  const int bj = loc % ts.nbj[r];
  const int bs = loc / ts.nbj[r];
  const Region& R = ts.r[r];
  t.r = r;
  t.k0 = R.k0 + bk * OZ;
  t.j0 = R.j0 + bj * OR;
  t.i0 = R.i0 + bs * ts.seg;
  t.i1 = min(t.i0 + ts.seg, R.i1);
  return true;
}

// CPML psi for one derivative: read old, write new only when `own`.
__device__ __forceinline__ void cpml_dd(const float* po, float* pn, long long pq, float b, float c, float& D,
                                        bool own) {
  const float v = b * po[pq] + c * D;
  if (own) pn[pq] = v;
  D += v;
}

}  // namespace

template <int STORE, bool GEN>
__global__ __launch_bounds__(TZ* TR) void k_torre(Grid g, Fields<STORE> fo, Fields<STORE> fn, Scales sc,
                                                  Cpml pmo, Cpml pmn, Materials mt, Ade ade, TorreSet ts) {
  __shared__ ClassEntry tbl[kMaxClasses];
  __shared__ float sE[2][3][ETILE];       // E tile of the current plane
  __shared__ float sH[4][3][TR][TZ];      // ring of new H planes
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * TZ + tx;
  for (int q = tid; q < mt.nclass; q += TZ * TR) tbl[q] = mt.table[q];
  TorreTile t;
  if (!torre_locate(ts, t)) return;  // block-uniform
  const Region& R = ts.r[t.r];

  // ---- this thread's column: H at (jh, kh); E owned when inside [1,14)x[1,62)
  const int jh = t.j0 - 1 + ty, kh = t.k0 - 1 + tx;
  const int jm = nbr(jh, g.ny, g.per[1]), km = nbr(kh, g.nz, g.per[2]);
  const bool hvalid = jm >= 0 && km >= 0;  // else H is the zero boundary value
  const long long hoff = hvalid ? (long long)jm * g.sy + km : 0;
  const bool eown = ty >= 1 && ty < TR - 2 && tx >= 1 && tx < TZ - 2 && jh >= R.j0 && jh < R.j1 && kh >= R.k0 &&
                    kh < R.k1 && jh >= t.j0 && jh < t.j0 + OR && kh >= t.k0 && kh < t.k0 + OZ;

  // ---- flat E-tile positions handled by this thread (up to two)
  long long foff[2];
  bool fval[2];
  int fpos[2];
  for (int u = 0; u < 2; u++) {
    const int p = tid + u * TZ * TR;
    fpos[u] = p < ETILE ? p : -1;
    fval[u] = false;
    foff[u] = 0;
    if (p < ETILE) {
      const int r = p / EC, c = p % EC;
      const int jj = nbr(t.j0 - 3 + r, g.ny, g.per[1]), kk = nbr(t.k0 - 3 + c, g.nz, g.per[2]);
      if (jj >= 0 && kk >= 0) {
        fval[u] = true;
        foff[u] = (long long)jj * g.sy + kk;
      }
    }
  }

  const auto* Exo = fo.E[0];
  const auto* Eyo = fo.E[1];
  const auto* Ezo = fo.E[2];
  const auto* Hxo = fo.H[0];
  const auto* Hyo = fo.H[1];
  const auto* Hzo = fo.H[2];
  const float iex = sc.ie[0], iey = sc.ie[1], iez = sc.ie[2];
  const float ihx = sc.ih[0], ihy = sc.ih[1], ihz = sc.ih[2];

  auto plane = [&](int i) -> long long {  // plane base under the x boundary rule (-1: zero)
    if (!GEN) return (long long)i * g.sx;
    const int ii = nbr(i, g.nx, g.per[0]);
    return ii < 0 ? -1LL : (long long)ii * g.sx;
  };

  // ---- x queues: Ey, Ez at planes p-2..p+1 (H stencil), Ex at p-2..p (E RMW)
  const int p0 = t.i0 - 1;  // first H plane
  float qy[4], qz[4], qx[3];
  for (int s = 0; s < 4; s++) {
    const long long pb = plane(p0 - 2 + s);
    const bool ok = hvalid && pb >= 0;
    qy[s] = ok ? ld(Eyo, pb + hoff, iey) : 0.f;
    qz[s] = ok ? ld(Ezo, pb + hoff, iez) : 0.f;
  }
  for (int s = 0; s < 2; s++) {  // qx[2] (plane p) comes from the tile centre
    const long long pb = plane(p0 - 2 + s);
    qx[s] = (hvalid && pb >= 0) ? ld(Exo, pb + hoff, iex) : 0.f;
  }
  qx[2] = 0.f;
  // new-H queue of this column: Hy, Hz at planes p-3..p (E stencil), Hx at p-2
  float hyq[4] = {0.f, 0.f, 0.f, 0.f}, hzq[4] = {0.f, 0.f, 0.f, 0.f}, hxq[3] = {0.f, 0.f, 0.f};

  // ---- prefetch registers for plane p: E tile values and old H
  float pe[2][3], ph[3];
  auto fetch = [&](int p) {
    const long long pb = plane(p);
    for (int u = 0; u < 2; u++) {
      const bool ok = fpos[u] >= 0 && fval[u] && pb >= 0;
      pe[u][0] = ok ? ld(Exo, pb + foff[u], iex) : 0.f;
      pe[u][1] = ok ? ld(Eyo, pb + foff[u], iey) : 0.f;
      pe[u][2] = ok ? ld(Ezo, pb + foff[u], iez) : 0.f;
    }
    const bool ok = hvalid && pb >= 0;
    ph[0] = ok ? ld(Hxo, pb + hoff, ihx) : 0.f;
    ph[1] = ok ? ld(Hyo, pb + hoff, ihy) : 0.f;
    ph[2] = ok ? ld(Hzo, pb + hoff, ihz) : 0.f;
  };
  fetch(p0);

  // per-column CPML state (z slab; x/y slabs only in GEN regions)
  const bool pm_on = pmo.on;
  const int zsH = (pm_on && hvalid) ? slab_index(km, g.nz, pmo.plo[2], pmo.phi[2]) : -1;
  const int ysH = (GEN && pm_on && hvalid) ? slab_index(jm, g.ny, pmo.plo[1], pmo.phi[1]) : -1;
  const bool design_col = mt.d_on && jh >= mt.d_j0 && jh < mt.d_j0 + mt.d_nj && kh >= mt.d_k0 &&
                          kh < mt.d_k0 + mt.d_nk;
  const float rd0 = g.rd[0], rd1 = g.rd[1], rd2 = g.rd[2], dt = g.dt;
  const int pend = t.i1 + 1;  // last H plane

  for (int p = p0; p <= pend; p++) {
    const int eb = p & 1, hb = p & 3;
    // 1) publish the E tile of plane p
    for (int u = 0; u < 2; u++) {
      if (fpos[u] < 0) continue;
      sE[eb][0][fpos[u]] = pe[u][0];
      sE[eb][1][fpos[u]] = pe[u][1];
      sE[eb][2][fpos[u]] = pe[u][2];
    }
    const float hx_old = ph[0], hy_old = ph[1], hz_old = ph[2];
    // 2) issue the loads for plane p+1 and the queue's plane p+2
    const long long pb = plane(p);
    float nqy = 0.f, nqz = 0.f;
    if (p < pend) {
      fetch(p + 1);
      const long long p2 = plane(p + 2);
      if (hvalid && p2 >= 0) {
        nqy = ld(Eyo, p2 + hoff, iey);
        nqz = ld(Ezo, p2 + hoff, iez);
      }
    }
    // E operands for plane m = p-2 (loaded early, used after the H half)
    const int m = p - 2;
    const bool edo = eown && m >= t.i0 && m < t.i1;
    int cls = 0;
    if (edo && mt.cls) cls = mt.cls[(long long)m * g.sx + hoff];
    lds_barrier();

    // 3) H at plane p (every thread; stored when owned)
    const int c0 = (ty + 2) * EC + (tx + 2);  // this column in the E tile
    const float* tEx = sE[eb][0];
    const float* tEy = sE[eb][1];
    const float* tEz = sE[eb][2];
    const float ex0 = tEx[c0];
    qx[2] = ex0;
    float hxn = 0.f, hyn = 0.f, hzn = 0.f;
    const bool hplane_ok = hvalid && pb >= 0;
    if (hplane_ok) {
      float dyEz = rd1 * d24(tEz[c0 - 2 * EC], tEz[c0 - EC], qz[2], tEz[c0 + EC]);
      float dyEx = rd1 * d24(tEx[c0 - 2 * EC], tEx[c0 - EC], ex0, tEx[c0 + EC]);
      float dzEy = rd2 * d24(tEy[c0 - 2], tEy[c0 - 1], qy[2], tEy[c0 + 1]);
      float dzEx = rd2 * d24(tEx[c0 - 2], tEx[c0 - 1], ex0, tEx[c0 + 1]);
      float dxEz = rd0 * d24(qz[0], qz[1], qz[2], qz[3]);
      float dxEy = rd0 * d24(qy[0], qy[1], qy[2], qy[3]);
      const int ip = GEN ? nbr(p, g.nx, g.per[0]) : p;
      const bool hown = eown && p >= t.i0 && p < t.i1;
      if (zsH >= 0) {  // z axis: slot 0 = Hx (D_z Ey), slot 1 = Hy (D_z Ex)
        const float bb = pmo.bH[2][km], cc = pmo.cH[2][km];
        const long long pq = psi_index(g, pmo, 2, zsH, ip, jm, km);
        cpml_dd(pmo.psiH[2][0], pmn.psiH[2][0], pq, bb, cc, dzEy, hown);
        cpml_dd(pmo.psiH[2][1], pmn.psiH[2][1], pq, bb, cc, dzEx, hown);
      }
      if (GEN && pm_on) {
        const int xs = slab_index(ip, g.nx, pmo.plo[0], pmo.phi[0]);
        if (xs >= 0) {  // x axis: slot 0 = Hy (D_x Ez), slot 1 = Hz (D_x Ey)
          const float bb = pmo.bH[0][ip], cc = pmo.cH[0][ip];
          const long long pq = psi_index(g, pmo, 0, xs, ip, jm, km);
          cpml_dd(pmo.psiH[0][0], pmn.psiH[0][0], pq, bb, cc, dxEz, hown);
          cpml_dd(pmo.psiH[0][1], pmn.psiH[0][1], pq, bb, cc, dxEy, hown);
        }
        if (ysH >= 0) {  // y axis: slot 0 = Hz (D_y Ex), slot 1 = Hx (D_y Ez)
          const float bb = pmo.bH[1][jm], cc = pmo.cH[1][jm];
          const long long pq = psi_index(g, pmo, 1, ysH, ip, jm, km);
          cpml_dd(pmo.psiH[1][0], pmn.psiH[1][0], pq, bb, cc, dyEx, hown);
          cpml_dd(pmo.psiH[1][1], pmn.psiH[1][1], pq, bb, cc, dyEz, hown);
        }
      }
      hxn = hx_old - dt * (dyEz - dzEy);
      hyn = hy_old - dt * (dzEx - dxEz);
      hzn = hz_old - dt * (dxEy - dyEx);
      if (hown) {
        const long long q = pb + hoff;
        st(fn.H[0], q, hxn, sc.sh[0]);
        st(fn.H[1], q, hyn, sc.sh[1]);
        st(fn.H[2], q, hzn, sc.sh[2]);
      }
    }
    // ring slot for plane p, and this column's queues
    sH[hb][0][ty][tx] = hxn;
    sH[hb][1][ty][tx] = hyn;
    sH[hb][2][ty][tx] = hzn;
    hyq[0] = hyq[1]; hyq[1] = hyq[2]; hyq[2] = hyq[3]; hyq[3] = hyn;
    hzq[0] = hzq[1]; hzq[1] = hzq[2]; hzq[2] = hzq[3]; hzq[3] = hzn;
    hxq[0] = hxq[1]; hxq[1] = hxq[2]; hxq[2] = hxn;

    // 4) E at plane m = p-2 from the ring slot written two planes ago
    if (edo) {
      const int mb = (p - 2) & 3;
      const float(*rHx)[TZ] = sH[mb][0];
      const float(*rHy)[TZ] = sH[mb][1];
      const float(*rHz)[TZ] = sH[mb][2];
      const float hx0 = hxq[0], hy0 = hyq[1], hz0 = hzq[1];
      float dyHz = rd1 * d24(rHz[ty - 1][tx], hz0, rHz[ty + 1][tx], rHz[ty + 2][tx]);
      float dyHx = rd1 * d24(rHx[ty - 1][tx], hx0, rHx[ty + 1][tx], rHx[ty + 2][tx]);
      float dzHy = rd2 * d24(rHy[ty][tx - 1], hy0, rHy[ty][tx + 1], rHy[ty][tx + 2]);
      float dzHx = rd2 * d24(rHx[ty][tx - 1], hx0, rHx[ty][tx + 1], rHx[ty][tx + 2]);
      float dxHz = rd0 * d24(hzq[0], hzq[1], hzq[2], hzq[3]);
      float dxHy = rd0 * d24(hyq[0], hyq[1], hyq[2], hyq[3]);
      if (pm_on) {
        const int zs = slab_index(km, g.nz, pmo.plo[2], pmo.phi[2]);
        if (zs >= 0) {  // z axis: slot 0 = Ex (D_z Hy), slot 1 = Ey (D_z Hx)
          const float bb = pmo.bE[2][km], cc = pmo.cE[2][km];
          const long long pq = psi_index(g, pmo, 2, zs, m, jm, km);
          cpml_dd(pmo.psiE[2][0], pmn.psiE[2][0], pq, bb, cc, dzHy, true);
          cpml_dd(pmo.psiE[2][1], pmn.psiE[2][1], pq, bb, cc, dzHx, true);
        }
        if (GEN) {
          const int xs = slab_index(m, g.nx, pmo.plo[0], pmo.phi[0]);
          if (xs >= 0) {  // x axis: slot 0 = Ey (D_x Hz), slot 1 = Ez (D_x Hy)
            const float bb = pmo.bE[0][m], cc = pmo.cE[0][m];
            const long long pq = psi_index(g, pmo, 0, xs, m, jm, km);
            cpml_dd(pmo.psiE[0][0], pmn.psiE[0][0], pq, bb, cc, dxHz, true);
            cpml_dd(pmo.psiE[0][1], pmn.psiE[0][1], pq, bb, cc, dxHy, true);
          }
          const int ys = slab_index(jm, g.ny, pmo.plo[1], pmo.phi[1]);
          if (ys >= 0) {  // y axis: slot 0 = Ez (D_y Hx), slot 1 = Ex (D_y Hz)
            const float bb = pmo.bE[1][jm], cc = pmo.cE[1][jm];
            const long long pq = psi_index(g, pmo, 1, ys, m, jm, km);
            cpml_dd(pmo.psiE[1][0], pmn.psiE[1][0], pq, bb, cc, dyHx, true);
            cpml_dd(pmo.psiE[1][1], pmn.psiE[1][1], pq, bb, cc, dyHz, true);
          }
        }
      }
      const float cx = dyHz - dzHy, cy = dzHx - dxHz, cz = dxHy - dyHx;
      const long long q = (long long)m * g.sx + hoff;
      const ClassEntry& ce = tbl[cls];
      float ie0 = ce.inv_eps[0], ie1 = ce.inv_eps[1], ie2 = ce.inv_eps[2];
      bool over = false;
      if (design_col && m >= mt.d_i0 && m < mt.d_i0 + mt.d_ni) {
        const float v = mt.d_inv_eps[((long long)(m - mt.d_i0) * mt.d_nj + (jh - mt.d_j0)) * mt.d_nk + (kh - mt.d_k0)];
        ie0 = ie1 = ie2 = v;
        over = true;
      }
      if (ce.special && !over) {  // finished by k_ade (old E -> new E)
        const long long bq = (long long)(m - ade.i0) * ade.bsx + (long long)(jh - ade.j0) * ade.bsy + (kh - ade.k0);
        ade.curl[0][bq] = cx;
        ade.curl[1][bq] = cy;
        ade.curl[2][bq] = cz;
      } else {
        st(fn.E[0], q, qx[0] + dt * ie0 * cx, sc.se[0]);
        st(fn.E[1], q, qy[0] + dt * ie1 * cy, sc.se[1]);
        st(fn.E[2], q, qz[0] + dt * ie2 * cz, sc.se[2]);
      }
    }
    // 5) advance the E queues (planes p-1 .. p+2) and Ex (p-1 .. p+1)
    qy[0] = qy[1]; qy[1] = qy[2]; qy[2] = qy[3]; qy[3] = nqy;
    qz[0] = qz[1]; qz[1] = qz[2]; qz[2] = qz[3]; qz[3] = nqz;
    qx[0] = qx[1]; qx[1] = qx[2];
  }
}

// ------------------------------------------------------------ launcher
static void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fprintf(stderr, "dtm: launch of %s failed: %s\n", what, hipGetErrorString(e));
}

TorreSet make_torre_set(const Region* regs, int nreg, int seg) {
  TorreSet s{};
  s.seg = seg;
  int acc = 0, m = 0;
  for (int t = 0; t < nreg && m < 6; t++) {
    const Region& r = regs[t];
    if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
    s.r[m] = r;
    s.nbk[m] = (r.k1 - r.k0 + OZ - 1) / OZ;
    s.nbj[m] = (r.j1 - r.j0 + OR - 1) / OR;
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

static int torre_seg(const Region* regs, int nreg) {
  long long cols = 0;
  int maxlen = 1;
  for (int t = 0; t < nreg; t++) {
    const Region& r = regs[t];
    if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
    cols += (long long)((r.k1 - r.k0 + OZ - 1) / OZ) * ((r.j1 - r.j0 + OR - 1) / OR);
    maxlen = std::max(maxlen, r.i1 - r.i0);
  }
  int seg = 64;
  while (seg > 8 && cols * ((maxlen + seg - 1) / seg) < 1024) seg /= 2;
  return seg;
}

template <int STORE>
static void torre_go(const Grid& g, const void* fo, const void* fn, const Scales& sc, const Cpml& pmo,
                     const Cpml& pmn, const Materials& mt, const Ade& ade, const Region* regs, int nreg,
                     const Region& main, hipStream_t st) {
  const Fields<STORE>& Fo = *static_cast<const Fields<STORE>*>(fo);
  const Fields<STORE>& Fn = *static_cast<const Fields<STORE>*>(fn);
  const dim3 blk(TZ, TR, 1);
  const int nm = main.i1 > main.i0 ? 1 : 0;
  TorreSet tm = make_torre_set(&main, nm, torre_seg(&main, nm));
  if (tm.total > 0) {
    hipLaunchKernelGGL((k_torre<STORE, false>), dim3(tm.per * NXCD), blk, 0, st, g, Fo, Fn, sc, pmo, pmn, mt, ade,
                       tm);
    check_launch("k_torre<main>");
  }
  TorreSet te = make_torre_set(regs, nreg, torre_seg(regs, nreg));
  if (te.total > 0) {
    hipLaunchKernelGGL((k_torre<STORE, true>), dim3(te.per * NXCD), blk, 0, st, g, Fo, Fn, sc, pmo, pmn, mt, ade,
                       te);
    check_launch("k_torre<edge>");
  }
}

void launch_torre(int store, const Grid& g, const void* fields_old, const void* fields_new, const Scales& sc,
                  const Cpml& pm_old, const Cpml& pm_new, const Materials& mt, const Ade& ade, const Region* regs,
                  int nreg, const Region& main, hipStream_t st) {
  if (store == 0) torre_go<0>(g, fields_old, fields_new, sc, pm_old, pm_new, mt, ade, regs, nreg, main, st);
  else torre_go<1>(g, fields_old, fields_new, sc, pm_old, pm_new, mt, ade, regs, nreg, main, st);
}

}  // namespace dtm
