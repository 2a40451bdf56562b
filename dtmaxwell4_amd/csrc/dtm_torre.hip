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
// Block = 64 lanes (z) x TR=16 rows (y).  On every plane p the block
// computes H on a 16 x 67 tile, rows j0-1 .. j0+14: lane tx at column
// k0+tx, and the first wave (row 0, which owns no E) also at the halo
// columns k0-1, k0+64, k0+65 of all 16 rows.  Two planes behind it computes
// E on the owned 13 x 64 tile (rows j0 .. j0+12, every lane).  Halo H values
// are recomputed by neighbouring blocks from the same old fields, so fields
// and CPML psi are double-buffered (read old, write new) and blocks need no
// synchronisation with each other.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>

#include "dtm_internal.h"

namespace dtm {

namespace {

constexpr int TZ = 64;            // lanes per row (one wavefront)
constexpr int TR = 16;            // thread rows
constexpr int OZ = TZ;            // owned z per tile (64)
constexpr int OR = TR - 3;        // owned rows per tile (13)
constexpr int HC = TZ + 3;        // H tile columns (67): k0-1 .. k0+65
constexpr int ER = TR + 3;        // E tile rows (19): j0-3 .. j0+15
constexpr int EC = TZ + 6;        // E tile cols (70): k0-3 .. k0+66
constexpr int ETILE = ER * EC;    // 1330
constexpr int NXCD = 8;
constexpr int kFoldPlanes = 264;  // planes of a block's fold table: segments of at most kFoldPlanes - 3

template <typename T> __device__ __forceinline__ float cv(T v, float inv);
template <> __device__ __forceinline__ float cv<float>(float v, float) { return v; }
template <> __device__ __forceinline__ float cv<__half>(__half v, float inv) { return __half2float(v) * inv; }
// Pointers read from the per-plane table are generic; casting them to the
// global address space keeps global_load/store (flat accesses would also
// count against the LDS wait counter that every barrier drains).
template <typename T> using GlobalT = __attribute__((address_space(1))) T;
template <typename T> __device__ __forceinline__ GlobalT<T>* gl(T* p) { return (GlobalT<T>*)p; }
// the table itself lives in the constant address space: scalar loads
using ColdK = __attribute__((address_space(4))) const TorreCold;
__device__ __forceinline__ void st(float* F, long long q, float v, float) { gl(F)[q] = v; }
__device__ __forceinline__ void st(__half* F, long long q, float v, float s) {
  gl(reinterpret_cast<unsigned short*>(F))[q] = __half_as_ushort(__float2half_rn(v * s));
}
// word at a position under the boundary rule (off < 0 or plane < 0: zero pad)
template <typename T>
__device__ __forceinline__ T ldz(const T* F, long long base, int off, unsigned zidx) {
  const unsigned idx = (base < 0 || off < 0) ? zidx : (unsigned)base + (unsigned)off;
  if constexpr (std::is_same<T, __half>::value)
    return __ushort_as_half(gl(reinterpret_cast<const unsigned short*>(F))[idx]);
  else
    return gl(F)[idx];
}

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
// psi index (psi arrays hold far fewer than 2^32 floats); P = slab cells of axis a
__device__ __forceinline__ unsigned psi_index(const Grid& g, unsigned P, int a, int s, int i, int j, int k) {
  if (a == 0) return ((unsigned)s * (unsigned)g.ny + (unsigned)j) * (unsigned)g.nz + (unsigned)k;
  if (a == 1) return ((unsigned)i * P + (unsigned)s) * (unsigned)g.nz + (unsigned)k;
  return ((unsigned)i * (unsigned)g.ny + (unsigned)j) * P + (unsigned)s;
}
// ring slot of plane p in an R-plane ring (p >= -8)
template <int R> __device__ __forceinline__ int slot(int p) { return (p + 4 * R) % R; }

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

// CPML psi for one derivative from the old value po: write new when `own`.
__device__ __forceinline__ void cpml_d(float po, float* pn, unsigned pq, float b, float c, float& D, bool own) {
  const float v = b * po + c * D;
  if (own) gl(pn)[pq] = v;
  D += v;
}


}  // namespace

// ============================================ fused step with x register queues
// The x stencils come from registers: every thread keeps, for its own (y, z)
// column, the four planes of E^n that the H update's x derivative reads and
// the four planes of H^{n+1/2} that the E update's reads (the wave-0 lanes
// that evaluate the halo H columns keep their own E queue).  LDS then holds
// only what the y/z stencils need: the E^n tile of the current plane (two
// buffers by plane parity) and the H tiles of the last four planes, 87 KB
// (LDS rings for the x stencils took 157 KB), and an x derivative costs no
// LDS traffic.  The plane
// loop is unrolled by four, so every queue slot and LDS buffer index is a
// compile-time constant (plane p of a tile lives in slot (p - p0) & 3).
namespace {
constexpr int NHALO = ETILE - TZ * TR;  // E tile positions outside the threads' own columns (306)

// Element of component c (Ex .. Hz) of a field set by a 32-bit byte offset
// from the set's base: one SGPR base for all six arrays, and the address is
// base + zero-extended offset (no 64-bit address arithmetic per access).
// The solver keeps a set below 4 GiB.
template <typename T>
__device__ __forceinline__ T ldc(const char* base, unsigned cb, int c, long long pb, int off, unsigned zidx) {
  const unsigned idx = (pb < 0 || off < 0) ? zidx : (unsigned)pb + (unsigned)off;
  const unsigned bo = (unsigned)c * cb + idx * (unsigned)sizeof(T);
  if constexpr (std::is_same<T, __half>::value)
    return __ushort_as_half(*reinterpret_cast<const GlobalT<unsigned short>*>(gl(base) + bo));
  else
    return *reinterpret_cast<const GlobalT<float>*>(gl(base) + bo);
}
template <typename T>
__device__ __forceinline__ void stc(char* base, unsigned cb, int c, unsigned idx, float v, float s) {
  const unsigned bo = (unsigned)c * cb + idx * (unsigned)sizeof(T);
  if constexpr (std::is_same<T, __half>::value)
    *reinterpret_cast<GlobalT<unsigned short>*>(gl(base) + bo) = __half_as_ushort(__float2half_rn(v * s));
  else
    *reinterpret_cast<GlobalT<float>*>(gl(base) + bo) = v;
}

template <typename T> struct PreT2 {
  T e[3];   // E^n at this thread's column, plane p+2
  T he[3];  // E^n at its halo tile position, plane p+1
  T h[3];   // H^{n-1/2} at its column, plane p+1
  T gh[3];  // wave-0 halo lanes: H^{n-1/2} at the halo column, plane p+1
  T ge[2];  // and E^n (Ey, Ez) there, plane p+2
  uint8_t cls;  // class byte of its E cell, plane p-1
};
}  // namespace

// TF: a plane-wave TF/SF source along x folded in, as in the two-pass sweeps:
// every H evaluation (own and halo columns) adds its weights times the H
// half-step's factors of plane p, every E update those of the E half-step at
// plane m.  fold: the H half-step's factor table, the E half-step's after it.
template <int STORE, int GEN, bool TF>
__global__ __launch_bounds__(TZ* TR) void k_torre2(Grid g, Scales sc, const uint8_t* clsg, const TorreCold* cold,
                                                   TorreSet ts, const float* fold) {
  using T = typename std::conditional<STORE == 0, float, __half>::type;
  constexpr bool GX = GEN & 1, GY = GEN & 2;
  __shared__ ClassEntry tbl[kMaxClasses];
  __shared__ float sE[2][3][ETILE];   // E^n tile of plane p, buffer (p - p0) & 1
  __shared__ float sH[4][3][TR][HC];  // H^{n+1/2} tiles, plane q in slot (q - p0) & 3
  // TF/SF fold factors of the block's planes p0 + l: the H half-step's at
  // [l][0-5], the E half-step's at [l][8-13]
  __shared__ __align__(16) float sF[TF ? kFoldPlanes : 1][16];
  // and each thread's packed weights (kept in LDS, not in registers held
  // across the plane loop): [0] own H column, [1] halo H column (wave 0) or own E cell
  __shared__ int sW[TF ? 2 : 1][TR][TZ];
  const int tx = threadIdx.x;
  const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one row per wavefront
  const int tid = ty * TZ + tx;
  const ColdK* C0 = (const ColdK*)cold;
  for (int q = tid; q < 4 * C0->mt.nclass; q += TZ * TR)
    reinterpret_cast<unsigned*>(tbl)[q] = gl(reinterpret_cast<const unsigned*>(C0->mt.table))[q];
  TorreTile t;
  if (!torre_locate(ts, t)) return;  // block-uniform
  const Region& R = ts.r[t.r];

  // own column: H at tile row ty (grid row j0-1+ty), tile column tx+1 (grid column k0+tx)
  const int jh = t.j0 - 1 + ty;
  const int jm = nbr(jh, g.ny, g.per[1]);
  const int khm = t.k0 + tx;
  const int kmm = nbr(khm, g.nz, g.per[2]);
  const bool hv_m = jm >= 0 && kmm >= 0;
  const int ho_m = hv_m ? jm * (int)g.sy + kmm : -1;
  const bool rown = ty >= 1 && ty < TR - 2 && jh >= R.j0 && jh < R.j1;
  const bool hown_c = rown && khm >= R.k0 && khm < R.k1;
  const bool eown = hown_c;
  const int emain = (ty + 2) * EC + (tx + 3);  // own column in the E tile
  // halo H columns k0-1, k0+64, k0+65 of all rows: wave 0, lane l -> row l/3
  const bool hal = ty == 0 && tx < 3 * TR;
  const int hr = hal ? tx / 3 : 0;
  const int chh = tx % 3 == 0 ? 0 : (tx % 3 == 1 ? HC - 2 : HC - 1);
  const int jmh = nbr(t.j0 - 1 + hr, g.ny, g.per[1]), kmh = nbr(t.k0 - 1 + chh, g.nz, g.per[2]);
  const bool hv_h = hal && jmh >= 0 && kmh >= 0;
  const int ho_h = hv_h ? jmh * (int)g.sy + kmh : -1;
  const int eh = (hr + 2) * EC + (chh + 2);  // halo column in the E tile
  // E tile positions outside the own columns: rows 0, 1, 18 (all 70 columns)
  // and columns 0-2, 67-69 of rows 2-17; thread tid < NHALO loads one
  int hpos = -1, hoff = -1;
  if (tid < NHALO) {
    int r, c;
    if (tid < 3 * EC) {
      r = tid / EC;
      if (r == 2) r = ER - 1;
      c = tid % EC;
    } else {
      const int q = tid - 3 * EC;
      r = 2 + q / 6;
      c = q % 6 < 3 ? q % 6 : EC - 6 + q % 6;
    }
    hpos = r * EC + c;
    const int jj = nbr(t.j0 - 3 + r, g.ny, g.per[1]), kk = nbr(t.k0 - 3 + c, g.nz, g.per[2]);
    if (jj >= 0 && kk >= 0) hoff = jj * (int)g.sy + kk;
  }
  const int ho_e = (eown && clsg) ? ho_m : -1;
  const uint8_t* clsp = clsg ? clsg : static_cast<const uint8_t*>(g.zero);
  const unsigned czidx = clsg ? g.zidx : 0u;
  const float ie[3] = {sc.ie[0], sc.ie[1], sc.ie[2]};
  const float ih[3] = {sc.ih[0], sc.ih[1], sc.ih[2]};

  auto plane = [&](int i) -> long long {
    if (!GX) return (long long)i * g.sx;
    const int ii = nbr(i, g.nx, g.per[0]);
    return ii < 0 ? -1LL : (long long)ii * g.sx;
  };
  // loads for the iteration after the one of plane p (clamped planes are harmless reloads)
  auto fetch = [&](PreT2<T>& S, int p, const ColdK* C) {
    const long long pe = plane(p + 2), ph = plane(p + 1), pc = plane(p - 1);
    const char* fb = C->fob;
    const unsigned cb = C->cbytes;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      S.e[c] = ldc<T>(fb, cb, c, pe, ho_m, g.zidx);
      S.he[c] = ldc<T>(fb, cb, c, ph, hoff, g.zidx);
      S.h[c] = ldc<T>(fb, cb, 3 + c, ph, ho_m, g.zidx);
    }
    if (ty == 0) {  // wave-uniform: the halo wave
#pragma unroll
      for (int c = 0; c < 3; c++) S.gh[c] = ldc<T>(fb, cb, 3 + c, ph, ho_h, g.zidx);
      S.ge[0] = ldc<T>(fb, cb, 1, pe, ho_h, g.zidx);
      S.ge[1] = ldc<T>(fb, cb, 2, pe, ho_h, g.zidx);
    }
    S.cls = ldz(clsp, clsg ? pc : 0, clsg ? ho_e : 0, czidx);
  };

  // CPML state per column / row
  const auto& zp = C0->zp;
  const bool pm_on = zp.on;
  const int xlo = zp.plo[0], xhi = zp.phi[0];
  const unsigned Px = zp.plo[0] + zp.phi[0], Py = zp.plo[1] + zp.phi[1], Pz = zp.plo[2] + zp.phi[2];
  const int zsm = (pm_on && hv_m) ? slab_index(kmm, g.nz, zp.plo[2], zp.phi[2]) : -1;
  const int zsh = (pm_on && hv_h) ? slab_index(kmh, g.nz, zp.plo[2], zp.phi[2]) : -1;
  const int ysl = (GY && pm_on && jm >= 0) ? slab_index(jm, g.ny, zp.plo[1], zp.phi[1]) : -1;
  const int yslh = (GY && pm_on && hv_h) ? slab_index(jmh, g.ny, zp.plo[1], zp.phi[1]) : -1;
  const auto& pc = C0->pmo;
  const float bzm = zsm >= 0 ? gl(pc.bH[2])[kmm] : 0.f, czm = zsm >= 0 ? gl(pc.cH[2])[kmm] : 0.f;
  const float bzh = zsh >= 0 ? gl(pc.bH[2])[kmh] : 0.f, czh = zsh >= 0 ? gl(pc.cH[2])[kmh] : 0.f;
  const float bze = zsm >= 0 ? gl(pc.bE[2])[kmm] : 0.f, cze = zsm >= 0 ? gl(pc.cE[2])[kmm] : 0.f;
  const float byh = ysl >= 0 ? gl(pc.bH[1])[jm] : 0.f, cyh = ysl >= 0 ? gl(pc.cH[1])[jm] : 0.f;
  const float byhh = yslh >= 0 ? gl(pc.bH[1])[jmh] : 0.f, cyhh = yslh >= 0 ? gl(pc.cH[1])[jmh] : 0.f;
  const float bye = ysl >= 0 ? gl(pc.bE[1])[jm] : 0.f, cye = ysl >= 0 ? gl(pc.cE[1])[jm] : 0.f;
  const auto& mc = C0->mt;
  const bool design_col = mc.d_on && jh >= mc.d_j0 && jh < mc.d_j0 + mc.d_nj && khm >= mc.d_k0 &&
                          khm < mc.d_k0 + mc.d_nk;
  const int di0 = mc.d_i0, di1 = mc.d_i0 + mc.d_ni;
  const float rd0 = g.rd[0], rd1 = g.rd[1], rd2 = g.rd[2], dt = g.dt;
  const int p0 = t.i0 - 1, pend = t.i1 + 1;  // H planes p0 .. pend
  // TF/SF fold: packed weights of the own H column, and of the halo H column
  // (wave 0, which owns no E) or the own E cell (zero away from the y and z faces)
  if constexpr (TF) {
    const float* foldE = fold + (long long)(g.nx + 1) * kFoldStride;
    const int* hd = reinterpret_cast<const int*>(fold);
    int tw2 = 0;
    if (ty == 0) tw2 = hv_h ? fold_pack(0, jmh, kmh, hd) : 0;
    else if (hv_m) tw2 = fold_pack(1, jm, kmm, hd);
    sW[0][ty][tx] = hv_m ? fold_pack(0, jm, kmm, hd) : 0;
    sW[TF ? 1 : 0][ty][tx] = tw2;
    // the factor tables of planes p0 .. pend (the launch keeps them within
    // kFoldPlanes); the first barrier of the plane loop orders these writes
    for (int q = tid; q < (pend - p0 + 1) * 16; q += TZ * TR) {
      const int l = q >> 4, c = q & 15;
      const int ii = min(max(GX ? nbr(p0 + l, g.nx, g.per[0]) : p0 + l, 0), g.nx - 1);
      sF[l][c] = (c & 7) < 6 ? (c < 8 ? fold : foldE)[kFoldStride * (1 + ii) + (c & 7)] : 0.f;
    }
  }

  // queues: qE[c][s] E^n of the own column, qG[c][s] Ey / Ez of the halo
  // column, qH[c][s] Hy / Hz of the own column; plane q in slot (q - p0) & 3
  float qE[3][4], qG[2][4], qH[2][4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    qE[0][s] = qE[1][s] = qE[2][s] = 0.f;
    qG[0][s] = qG[1][s] = qH[0][s] = qH[1][s] = 0.f;
  }
#pragma unroll
  for (int d = -2; d <= 0; d++) {  // planes p0-2 .. p0
    const long long pb = plane(p0 + d);
    const int s = d & 3;
#pragma unroll
    for (int c = 0; c < 3; c++) qE[c][s] = cv(ldc<T>(C0->fob, C0->cbytes, c, pb, ho_m, g.zidx), ie[c]);
    if (ty == 0) {
      qG[0][s] = cv(ldc<T>(C0->fob, C0->cbytes, 1, pb, ho_h, g.zidx), ie[1]);
      qG[1][s] = cv(ldc<T>(C0->fob, C0->cbytes, 2, pb, ho_h, g.zidx), ie[2]);
    }
  }
  PreT2<T> S;
  S = PreT2<T>{};
  fetch(S, min(p0 - 1, pend), C0);  // the loads of the iteration before p0

  // H from the E tile (y/z) and the x queue (planes p-2 .. p+1 in ex[0..3])
  auto h_new = [&](const ColdK* C, const float* tE, int e, const float* exy, const float* exz, int ip, int jr,
                   int km, int zs, float bz, float cz, int ys, float by, float cy, bool own, float hx0, float hy0,
                   float hz0, float pz0, float pz1, float py0, float py1, int tw, const float* fu, float& hx,
                   float& hy, float& hz) {
    const float* tEx = tE;
    const float* tEy = tE + ETILE;
    const float* tEz = tE + 2 * ETILE;
    const float ex0 = tEx[e];
    float dyEz = rd1 * d24(tEz[e - 2 * EC], tEz[e - EC], tEz[e], tEz[e + EC]);
    float dyEx = rd1 * d24(tEx[e - 2 * EC], tEx[e - EC], ex0, tEx[e + EC]);
    float dzEy = rd2 * d24(tEy[e - 2], tEy[e - 1], tEy[e], tEy[e + 1]);
    float dzEx = rd2 * d24(tEx[e - 2], tEx[e - 1], ex0, tEx[e + 1]);
    float dxEz = rd0 * d24(exz[0], exz[1], exz[2], exz[3]);
    float dxEy = rd0 * d24(exy[0], exy[1], exy[2], exy[3]);
    if constexpr (TF) {
      if (tw) {
        dyEz += fold_f(tw, 0) * fu[0];   // Hx, y faces
        dzEy += fold_f(tw, 6) * fu[1];   // Hx, z faces
        dzEx += fold_f(tw, 12) * fu[2];  // Hy, z faces
        dxEz += fold_b(tw, 24) * fu[3];  // Hy, x faces
        dxEy += fold_b(tw, 25) * fu[4];  // Hz, x faces
        dyEx += fold_f(tw, 18) * fu[5];  // Hz, y faces
      }
    }
    if (zs >= 0) {
      const unsigned pq = psi_index(g, Pz, 2, zs, ip, jr, km);
      cpml_d(pz0, C->zp.hN[0], pq, bz, cz, dzEy, own);
      cpml_d(pz1, C->zp.hN[1], pq, bz, cz, dzEx, own);
    }
    if (GX && pm_on) {
      const int xs = slab_index(ip, g.nx, xlo, xhi);
      if (xs >= 0) {
        const float bb = gl(C->pmo.bH[0])[ip], cc = gl(C->pmo.cH[0])[ip];
        const unsigned pq = psi_index(g, Px, 0, xs, ip, jr, km);
        cpml_d(gl(C->pmo.psiH[0][0])[pq], C->pmn.psiH[0][0], pq, bb, cc, dxEz, own);
        cpml_d(gl(C->pmo.psiH[0][1])[pq], C->pmn.psiH[0][1], pq, bb, cc, dxEy, own);
      }
    }
    if (GY && ys >= 0) {
      const unsigned pq = psi_index(g, Py, 1, ys, ip, jr, km);
      cpml_d(py0, C->pmn.psiH[1][0], pq, by, cy, dyEx, own);
      cpml_d(py1, C->pmn.psiH[1][1], pq, by, cy, dyEz, own);
    }
    hx = hx0 - dt * (dyEz - dzEy);
    hy = hy0 - dt * (dzEx - dxEz);
    hz = hz0 - dt * (dxEy - dyEx);
  };

  auto step = [&](auto U, int p) {
    constexpr int u = decltype(U)::value;
    constexpr int b = u & 1, s0 = u, s1 = (u + 1) & 3, s2 = (u + 2) & 3, s3 = (u + 3) & 3;
    // s0: plane p, s1: p+1 (and p-3), s2: p-2, s3: p-1
    const int m = p - 2;
    const TorreCold* Cg = cold;
    asm volatile("" : "+s"(Cg));
    const ColdK* C = (const ColdK*)Cg;
    // 1) take the loads issued last iteration
#pragma unroll
    for (int c = 0; c < 3; c++) qE[c][s1] = cv(S.e[c], ie[c]);
    float hE[3], oH[3], oG[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      hE[c] = cv(S.he[c], ie[c]);
      oH[c] = cv(S.h[c], ih[c]);
      oG[c] = cv(S.gh[c], ih[c]);
    }
    qG[0][s1] = cv(S.ge[0], ie[1]);
    qG[1][s1] = cv(S.ge[1], ie[2]);
    const uint8_t cls = S.cls;
    // 2) publish the E^n tile of plane p
    float* tE = &sE[b][0][0];
#pragma unroll
    for (int c = 0; c < 3; c++) tE[c * ETILE + emain] = qE[c][s0];
    if (hpos >= 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) tE[c * ETILE + hpos] = hE[c];
    }
    // the fold factors of H plane p and E plane m (read only by lanes next to a face)
    const float* fuH = sF[TF ? p - p0 : 0];
    const float* fuE = sF[TF ? max(m - p0, 0) : 0] + 8;
    // 3) psi and design operands of this plane, then the next iteration's loads
    const long long pb = plane(p);
    const int ip = GX ? nbr(p, g.nx, g.per[0]) : p;
    const bool hplane = pb >= 0;
    const int im = GX ? nbr(m, g.nx, g.per[0]) : m;
    const bool edo = eown && m >= t.i0 && m < t.i1;
    float pzm0 = 0.f, pzm1 = 0.f, pzh0 = 0.f, pzh1 = 0.f, pym0 = 0.f, pym1 = 0.f, pyh0 = 0.f, pyh1 = 0.f;
    float pze0 = 0.f, pze1 = 0.f, pye0 = 0.f, pye1 = 0.f, dv = 0.f;
    if (hplane && zsm >= 0) {
      const unsigned pq = psi_index(g, Pz, 2, zsm, ip, jm, kmm);
      pzm0 = gl(C->zp.hO[0])[pq];
      pzm1 = gl(C->zp.hO[1])[pq];
    }
    if (hplane && zsh >= 0) {
      const unsigned pq = psi_index(g, Pz, 2, zsh, ip, jmh, kmh);
      pzh0 = gl(C->zp.hO[0])[pq];
      pzh1 = gl(C->zp.hO[1])[pq];
    }
    if (GY && hplane && ysl >= 0 && hv_m) {
      const unsigned pq = psi_index(g, Py, 1, ysl, ip, jm, kmm);
      pym0 = gl(C->pmo.psiH[1][0])[pq];
      pym1 = gl(C->pmo.psiH[1][1])[pq];
    }
    if (GY && hplane && yslh >= 0) {
      const unsigned qh = psi_index(g, Py, 1, yslh, ip, jmh, kmh);
      pyh0 = gl(C->pmo.psiH[1][0])[qh];
      pyh1 = gl(C->pmo.psiH[1][1])[qh];
    }
    if (edo && zsm >= 0) {
      const unsigned pq = psi_index(g, Pz, 2, zsm, im, jm, kmm);
      pze0 = gl(C->zp.eO[0])[pq];
      pze1 = gl(C->zp.eO[1])[pq];
    }
    if (GY && edo && ysl >= 0) {
      const unsigned pq = psi_index(g, Py, 1, ysl, im, jm, kmm);
      pye0 = gl(C->pmo.psiE[1][0])[pq];
      pye1 = gl(C->pmo.psiE[1][1])[pq];
    }
    const bool over = edo && design_col && m >= di0 && m < di1;
    if (over) {
      const auto& M = C->mt;
      dv = gl(M.d_inv_eps)[((long long)(m - M.d_i0) * M.d_nj + (jh - M.d_j0)) * M.d_nk + (khm - M.d_k0)];
    }
    fetch(S, min(p, pend), C);
    lds_barrier();

    // 4) H at plane p, own column, then the halo columns (wave 0)
    const bool hown = hown_c && p >= t.i0 && p < t.i1;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    if (hplane && hv_m) {
      const float exy[4] = {qE[1][s2], qE[1][s3], qE[1][s0], qE[1][s1]};
      const float exz[4] = {qE[2][s2], qE[2][s3], qE[2][s0], qE[2][s1]};
      h_new(C, tE, emain, exy, exz, ip, jm, kmm, zsm, bzm, czm, ysl, byh, cyh, hown, oH[0], oH[1], oH[2], pzm0,
            pzm1, pym0, pym1, TF ? sW[0][ty][tx] : 0, fuH, hx, hy, hz);
    }
    sH[s0][0][ty][tx + 1] = hx;
    sH[s0][1][ty][tx + 1] = hy;
    sH[s0][2][ty][tx + 1] = hz;
    qH[0][s0] = hy;
    qH[1][s0] = hz;
    if (hown) {
      const unsigned q = (unsigned)pb + (unsigned)ho_m;
      stc<T>(C->fnb, C->cbytes, 3, q, hx, sc.sh[0]);
      stc<T>(C->fnb, C->cbytes, 4, q, hy, sc.sh[1]);
      stc<T>(C->fnb, C->cbytes, 5, q, hz, sc.sh[2]);
    }
    if (ty == 0) {
      float gx = 0.f, gy = 0.f, gz = 0.f;
      if (hplane && hv_h) {
        const float gey[4] = {qG[0][s2], qG[0][s3], qG[0][s0], qG[0][s1]};
        const float gez[4] = {qG[1][s2], qG[1][s3], qG[1][s0], qG[1][s1]};
        h_new(C, tE, eh, gey, gez, ip, jmh, kmh, zsh, bzh, czh, yslh, byhh, cyhh, false, oG[0], oG[1], oG[2], pzh0,
              pzh1, pyh0, pyh1, TF ? sW[TF ? 1 : 0][ty][tx] : 0, fuH, gx, gy, gz);
      }
      if (hal) {
        sH[s0][0][hr][chh] = gx;
        sH[s0][1][hr][chh] = gy;
        sH[s0][2][hr][chh] = gz;
      }
    }

    // 5) E at plane m = p-2 (owned cells): y/z from the H tiles of plane m,
    // x from the H queue (planes m-1 .. m+2 in slots s1, s2, s3, s0)
    if (edo) {
      const int c = tx + 1;
      const float(*h0)[TR][HC] = sH[s2];
      const float hx0 = h0[0][ty][c], hy0 = h0[1][ty][c], hz0 = h0[2][ty][c];
      float dyHz = rd1 * d24(h0[2][ty - 1][c], hz0, h0[2][ty + 1][c], h0[2][ty + 2][c]);
      float dyHx = rd1 * d24(h0[0][ty - 1][c], hx0, h0[0][ty + 1][c], h0[0][ty + 2][c]);
      float dzHy = rd2 * d24(h0[1][ty][c - 1], hy0, h0[1][ty][c + 1], h0[1][ty][c + 2]);
      float dzHx = rd2 * d24(h0[0][ty][c - 1], hx0, h0[0][ty][c + 1], h0[0][ty][c + 2]);
      float dxHz = rd0 * d24(qH[1][s1], qH[1][s2], qH[1][s3], qH[1][s0]);
      float dxHy = rd0 * d24(qH[0][s1], qH[0][s2], qH[0][s3], qH[0][s0]);
      if constexpr (TF) {
        const int tw2 = sW[TF ? 1 : 0][ty][tx];
        if (tw2) {
          dyHz += fold_f(tw2, 0) * fuE[0];   // Ex, y faces
          dzHy += fold_f(tw2, 6) * fuE[1];   // Ex, z faces
          dzHx += fold_f(tw2, 12) * fuE[2];  // Ey, z faces
          dxHz += fold_b(tw2, 24) * fuE[3];  // Ey, x faces
          dxHy += fold_b(tw2, 25) * fuE[4];  // Ez, x faces
          dyHx += fold_f(tw2, 18) * fuE[5];  // Ez, y faces
        }
      }
      if (zsm >= 0) {
        const unsigned pq = psi_index(g, Pz, 2, zsm, im, jm, kmm);
        cpml_d(pze0, C->zp.eN[0], pq, bze, cze, dzHy, true);
        cpml_d(pze1, C->zp.eN[1], pq, bze, cze, dzHx, true);
      }
      if (GX && pm_on) {
        const int xs = slab_index(im, g.nx, xlo, xhi);
        if (xs >= 0) {
          const float bb = gl(C->pmo.bE[0])[im], cc = gl(C->pmo.cE[0])[im];
          const unsigned pq = psi_index(g, Px, 0, xs, im, jm, kmm);
          cpml_d(gl(C->pmo.psiE[0][0])[pq], C->pmn.psiE[0][0], pq, bb, cc, dxHz, true);
          cpml_d(gl(C->pmo.psiE[0][1])[pq], C->pmn.psiE[0][1], pq, bb, cc, dxHy, true);
        }
      }
      if (GY && ysl >= 0) {
        const unsigned pq = psi_index(g, Py, 1, ysl, im, jm, kmm);
        cpml_d(pye0, C->pmn.psiE[1][0], pq, bye, cye, dyHx, true);
        cpml_d(pye1, C->pmn.psiE[1][1], pq, bye, cye, dyHz, true);
      }
      const float cx = dyHz - dzHy, cy = dzHx - dxHz, cz = dxHy - dyHx;
      const long long q = (long long)im * g.sx + ho_m;
      const ClassEntry& ce = tbl[cls];
      float ie0 = ce.inv_eps[0], ie1 = ce.inv_eps[1], ie2 = ce.inv_eps[2];
      if (over) ie0 = ie1 = ie2 = dv;
      if (ce.special && !over) {  // finished by k_ade
        const auto& A = C->ade;
        const long long bq = (long long)(m - A.i0) * A.bsx + (long long)(jh - A.j0) * A.bsy + (khm - A.k0);
        gl(A.curl[0])[bq] = cx;
        gl(A.curl[1])[bq] = cy;
        gl(A.curl[2])[bq] = cz;
      } else {
        stc<T>(C->fnb, C->cbytes, 0, (unsigned)q, qE[0][s2] + dt * ie0 * cx, sc.se[0]);
        stc<T>(C->fnb, C->cbytes, 1, (unsigned)q, qE[1][s2] + dt * ie1 * cy, sc.se[1]);
        stc<T>(C->fnb, C->cbytes, 2, (unsigned)q, qE[2][s2] + dt * ie2 * cz, sc.se[2]);
      }
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  int p = p0;
  for (; p + 3 <= pend; p += 4) {
    step(I0{}, p);
    step(I1{}, p + 1);
    step(I2{}, p + 2);
    step(I3{}, p + 3);
  }
  if (p <= pend) step(I0{}, p);
  if (p + 1 <= pend) step(I1{}, p + 1);
  if (p + 2 <= pend) step(I2{}, p + 2);
}

// ------------------------------------------------------------ launcher
static void check_launch(const char* what) { note_launch(what); }

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

// Segment length: a block computes seg + 3 H planes (one halo plane below,
// two above) after a warm-up of about kWarm planes, and one launch runs in
// ceil(blocks / slots) rounds; minimise rounds * (seg + 3 + kWarm) over
// segments of at most maxseg planes.
static int torre_seg(const Region* regs, int nreg, int slots, int maxseg) {
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
    if (nseg > 1 && seg == (maxlen + nseg - 2) / (nseg - 1)) continue;
    if (seg > maxseg) continue;
    long long blocks = 0;
    for (int t = 0; t < nreg; t++) {
      const Region& r = regs[t];
      if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) continue;
      blocks += (long long)((r.k1 - r.k0 + OZ - 1) / OZ) * ((r.j1 - r.j0 + OR - 1) / OR) * ((r.i1 - r.i0 + seg - 1) / seg);
    }
    const long long cost = ((blocks + slots - 1) / slots) * (seg + 3 + kWarm);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = seg;
    }
    if (seg <= 4) break;
  }
  return best;
}

template <typename K>
static int torre_slots(K kernel) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
    cus = p.multiProcessorCount;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, TZ * TR, 0) != hipSuccess || per < 1) per = 1;
  return per * cus;
}

template <int STORE>
static void torre_go(const Grid& g, const Scales& sc, const Materials& mt, const TorreCold* cold,
                     const Region* regs, int nreg, hipStream_t st, const float* fold) {
  const dim3 blk(TZ, TR, 1);
  static const int slots = torre_slots(k_torre2<STORE, 3, false>);
  // a folded source keeps each block's factor table within kFoldPlanes
  TorreSet te = make_torre_set(regs, nreg, torre_seg(regs, nreg, slots, fold ? kFoldPlanes - 3 : 1 << 30));
  if (te.total <= 0) return;
  if (fold)
    hipLaunchKernelGGL((k_torre2<STORE, 3, true>), dim3(te.per * NXCD), blk, 0, st, g, sc, mt.cls, cold, te, fold);
  else
    hipLaunchKernelGGL((k_torre2<STORE, 3, false>), dim3(te.per * NXCD), blk, 0, st, g, sc, mt.cls, cold, te, fold);
  check_launch("k_torre2");
}

template <int STORE>
static void cold_fields(TorreCold& c, const void* fo, const void* fn) {
  const Fields<STORE>& Fo = *static_cast<const Fields<STORE>*>(fo);
  const Fields<STORE>& Fn = *static_cast<const Fields<STORE>*>(fn);
  for (int q = 0; q < 3; q++) {
    c.fo[q] = Fo.E[q];
    c.fo[3 + q] = Fo.H[q];
    c.fn[q] = Fn.E[q];
    c.fn[3 + q] = Fn.H[q];
  }
  // the solver allocates each field set as one block, Ex .. Hz at equal strides
  c.fob = reinterpret_cast<const char*>(Fo.E[0]);
  c.fnb = reinterpret_cast<char*>(Fn.E[0]);
  c.cbytes = (unsigned)(reinterpret_cast<const char*>(Fo.E[1]) - reinterpret_cast<const char*>(Fo.E[0]));
}

TorreCold make_torre_cold(const void* fields_old, const void* fields_new, int store, const Cpml& pm_old,
                          const Cpml& pm_new, const Materials& mt, const Ade& ade) {
  TorreCold c;
  memset(&c, 0, sizeof(c));  // padding too: callers compare tables bytewise
  if (store == 0) cold_fields<0>(c, fields_old, fields_new);
  else cold_fields<1>(c, fields_old, fields_new);
  c.zp.on = pm_old.on;
  for (int a = 0; a < 3; a++) {
    c.zp.plo[a] = pm_old.plo[a];
    c.zp.phi[a] = pm_old.phi[a];
  }
  for (int sl = 0; sl < 2; sl++) {
    c.zp.hO[sl] = pm_old.psiH[2][sl];
    c.zp.hN[sl] = pm_new.psiH[2][sl];
    c.zp.eO[sl] = pm_old.psiE[2][sl];
    c.zp.eN[sl] = pm_new.psiE[2][sl];
  }
  memcpy(&c.pmo, &pm_old, sizeof(Cpml));
  memcpy(&c.pmn, &pm_new, sizeof(Cpml));
  memcpy(&c.mt, &mt, sizeof(Materials));
  memcpy(&c.ade, &ade, sizeof(Ade));
  return c;
}

void launch_torre(int store, const Grid& g, const Scales& sc, const Materials& mt, const TorreCold* cold,
                  const Region* regs, int nreg, hipStream_t st, const float* fold) {
  if (store == 0) torre_go<0>(g, sc, mt, cold, regs, nreg, st, fold);
  else torre_go<1>(g, sc, mt, cold, regs, nreg, st, fold);
}

}  // namespace dtm
