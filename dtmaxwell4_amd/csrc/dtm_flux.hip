// dtm_flux.hip — FP64 Poynting-flux sums and box energies on the device.
//
// Flux boxes.  A box of cells [i0,i1) x [j0,j1) x [k0,k1) has six faces on
// the H planes a = i0, i1 (x), j0, j1 (y), k0, k1 (z).  Through the face
// normal to axis a (tangential axes b = a+1, c = a+2, mod 3)
//     S_a = E_b H_c - E_c H_b.
// On the lattice of include/dtmaxwell4.h, E_b and H_c share their b and c
// coordinates (even along b, odd along c) and E_c and H_b share theirs (odd
// along b, even along c); H_b and H_c lie on the face plane and E_b, E_c half
// a cell off it on either side.  So per face node
//     term 1: avg_a(E_b) H_c on nodes (b even: trapezoid over [b0, b1],
//                                       c odd:  midpoint over [c0, c1))
//     term 2: avg_a(E_c) H_b on nodes (b odd:  midpoint over [b0, b1),
//                                       c even: trapezoid over [c0, c1])
// with avg_a the mean of the cells on both sides of the face plane: a
// second-order surface integral of the continuous flux.  In time the step
// n -> n+1 gives S^{n+1/2} from H^{n+1/2} and (E^n + E^{n+1})/2: k_flux_snap
// stores the face-averaged E^n before the step, k_flux combines it with the
// fields after the step.  Each face is reduced by one block in a fixed
// order, so the FP64 sums are reproducible bit for bit.
//
// No counterpart in the reference, which writes only point sensors
// (sensor.h); the FP64 flux sums are part of this path's precision contract.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <mutex>

#include "dtm_internal.h"

namespace dtm {

namespace {

__device__ __forceinline__ float ldf(const float* F, long long q, float) { return F[q]; }
__device__ __forceinline__ float ldf(const __half* F, long long q, float inv) { return __half2float(F[q]) * inv; }

// Face node -> (term, position) of one face; see FluxFace.
struct FaceNode {
  int term;        // 0: E_b H_c, 1: E_c H_b
  int pa, pb, pc;  // cell coordinates along a, b, c of the H value
  double w;        // trapezoid weight (1 or 1/2 per even axis end)
};

__device__ __forceinline__ FaceNode face_node(const FluxFace& F, int n) {
  FaceNode r;
  const int nb = F.b1 - F.b0, nc = F.c1 - F.c0;
  const int n1 = (nb + 1) * nc;  // term-1 nodes: b in [b0,b1], c in [c0,c1)
  r.pa = F.pa;
  if (n < n1) {
    r.term = 0;
    const int ib = n / nc, ic = n % nc;
    r.pb = F.b0 + ib;
    r.pc = F.c0 + ic;
    r.w = (ib == 0 || ib == nb) ? 0.5 : 1.0;
  } else {
    n -= n1;  // term-2 nodes: b in [b0,b1), c in [c0,c1]
    r.term = 1;
    const int ib = n / (nc + 1), ic = n % (nc + 1);
    r.pb = F.b0 + ib;
    r.pc = F.c0 + ic;
    r.w = (ic == 0 || ic == nc) ? 0.5 : 1.0;
  }
  return r;
}

__device__ __forceinline__ long long cell_q(const Grid& g, int a, int pa, int pb, int pc) {
  int v[3];
  v[a] = pa;
  v[(a + 1) % 3] = pb;
  v[(a + 2) % 3] = pc;
  return (long long)v[0] * g.sx + (long long)v[1] * g.sy + v[2];
}

template <typename T>
__device__ __forceinline__ float face_avg_e(const Grid& g, const T* E, float inv, int a, const FaceNode& nd) {
  const long long q1 = cell_q(g, a, nd.pa, nd.pb, nd.pc);
  const long long q0 = q1 - (a == 0 ? g.sx : (a == 1 ? g.sy : 1));
  return 0.5f * (ldf(E, q0, inv) + ldf(E, q1, inv));
}

// Face-averaged tangential E^n of every face node (before the step).
template <int STORE>
__global__ __launch_bounds__(256) void k_flux_snap(Grid g, Fields<STORE> f, Scales sc, FluxSet fs) {
  const int face = blockIdx.y;
  const FluxFace& F = fs.face[face];
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < F.nodes; n += gridDim.x * blockDim.x) {
    const FaceNode nd = face_node(F, n);
    const int ce = nd.term == 0 ? (F.a + 1) % 3 : (F.a + 2) % 3;  // E_b or E_c
    fs.snap[F.snap0 + n] = face_avg_e(g, f.E[ce], sc.ie[ce], F.a, nd);
  }
}

// Deterministic block sum of 256 doubles (fixed tree order).
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// One block per face: outward flux S^{n+1/2}.n dA summed in FP64; written to
// the series row and added (times dt) to the face's running total.
template <int STORE>
__global__ __launch_bounds__(256) void k_flux(Grid g, Fields<STORE> f, Scales sc, FluxSet fs, int row, double dt) {
  __shared__ double red[256];
  const int face = blockIdx.x;
  const FluxFace& F = fs.face[face];
  const int b = (F.a + 1) % 3, c = (F.a + 2) % 3;
  double acc = 0.0;
  for (int n = threadIdx.x; n < F.nodes; n += blockDim.x) {
    const FaceNode nd = face_node(F, n);
    const int ce = nd.term == 0 ? b : c;  // E_b with H_c, E_c with H_b
    const int ch = nd.term == 0 ? c : b;
    const float e1 = face_avg_e(g, f.E[ce], sc.ie[ce], F.a, nd);
    const float e = 0.5f * (fs.snap[F.snap0 + n] + e1);
    const float h = ldf(f.H[ch], cell_q(g, F.a, nd.pa, nd.pb, nd.pc), sc.ih[ch]);
    const double term = (double)e * (double)h * nd.w;
    acc += nd.term == 0 ? term : -term;
  }
  const double s = block_sum(acc, red) * F.area * F.sign;
  if (threadIdx.x == 0) {
    fs.series[(long long)row * fs.nface + face] = s;
    fs.total[face] += s * dt;
  }
}

// Box energy, stage 1: one block per x plane of the box, the plane's
// 1/2 (eps E^2 + H^2) summed in FP64 into part[plane].
template <int STORE>
__global__ __launch_bounds__(256) void k_energy_planes(Grid g, Fields<STORE> f, Scales sc, Materials mt, Region bx,
                                                       double* part) {
  __shared__ double red[256];
  const int i = bx.i0 + blockIdx.x;
  const int nj = bx.j1 - bx.j0, nk = bx.k1 - bx.k0;
  double acc = 0.0;
  for (int t = threadIdx.x; t < nj * nk; t += blockDim.x) {
    const int j = bx.j0 + t / nk, k = bx.k0 + t % nk;
    const long long q = i * g.sx + j * g.sy + k;
    const ClassEntry& ce = mt.table[mt.cls ? mt.cls[q] : 0];
    float ie[3] = {ce.inv_eps[0], ce.inv_eps[1], ce.inv_eps[2]};
    if (mt.d_on) {
      const int di = i - mt.d_i0, dj = j - mt.d_j0, dk = k - mt.d_k0;
      if (di >= 0 && di < mt.d_ni && dj >= 0 && dj < mt.d_nj && dk >= 0 && dk < mt.d_nk)
        ie[0] = ie[1] = ie[2] = mt.d_inv_eps[((long long)di * mt.d_nj + dj) * mt.d_nk + dk];
    }
    for (int c = 0; c < 3; c++) {
      const double e = ldf(f.E[c], q, sc.ie[c]), h = ldf(f.H[c], q, sc.ih[c]);
      acc += e * e / ie[c] + h * h;
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = 0.5 * s;
}

// Stage 2: one block sums the plane partials in a fixed order.
__global__ __launch_bounds__(256) void k_energy_final(const double* part, int n, double dv, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) acc += part[t];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) *out = s * dv;
}

}  // namespace

// ------------------------------------------- box energy without allocation
// Scratch of dtm_energy_box_noalloc: the plane partials of one box (at most
// kStaticPlanes x planes) and the result in the last slot.  Every device the
// library runs on holds its own copy until the process ends; calls that
// overlap on one device share it and race.  Nothing in it must start at
// zero: each call overwrites the partials it sums and the result slot.
constexpr int kStaticPlanes = 8192;
static __device__ __align__(256) double scratch[kStaticPlanes + 1];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int static_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != dev) return -1;  // the device must be current
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return -1;
  g_scratch_ptr[dev] = static_cast<double*>(p);  // this device's copy; host code never uses `scratch` itself
  return 0;
}

int static_scratch_planes() { return kStaticPlanes; }

int energy_box_static(int dev, int store, const Grid& g, const void* fields, const Scales& sc, const Materials& mt,
                      const Region& bx, double dv, double* host_out, hipStream_t st) {
  if (bx.i1 - bx.i0 > kStaticPlanes) return -2;
  if (static_scratch_init(dev)) return -1;
  double* ptr = g_scratch_ptr[dev];
  launch_energy_box(store, g, fields, sc, mt, bx, ptr, dv, ptr + kStaticPlanes, st);
  if (hipMemcpyAsync(host_out, ptr + kStaticPlanes, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return -1;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

void launch_flux_snap(int store, const Grid& g, const void* fields, const Scales& sc, const FluxSet& fs,
                      hipStream_t st) {
  if (fs.nface <= 0) return;
  int most = 1;
  for (int q = 0; q < fs.nface; q++) most = fs.face[q].nodes > most ? fs.face[q].nodes : most;
  const dim3 grid((unsigned)((most + 255) / 256 < 1024 ? (most + 255) / 256 : 1024), (unsigned)fs.nface);
  if (store == 0)
    hipLaunchKernelGGL(k_flux_snap<0>, grid, dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fields), sc, fs);
  else
    hipLaunchKernelGGL(k_flux_snap<1>, grid, dim3(256), 0, st, g, *static_cast<const Fields<1>*>(fields), sc, fs);
  note_launch("k_flux_snap");
}

void launch_flux(int store, const Grid& g, const void* fields, const Scales& sc, const FluxSet& fs, int row,
                 double dt, hipStream_t st) {
  if (fs.nface <= 0) return;
  if (store == 0)
    hipLaunchKernelGGL(k_flux<0>, dim3(fs.nface), dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fields), sc,
                       fs, row, dt);
  else
    hipLaunchKernelGGL(k_flux<1>, dim3(fs.nface), dim3(256), 0, st, g, *static_cast<const Fields<1>*>(fields), sc,
                       fs, row, dt);
  note_launch("k_flux");
}

void launch_energy_box(int store, const Grid& g, const void* fields, const Scales& sc, const Materials& mt,
                       const Region& bx, double* part, double dv, double* out, hipStream_t st) {
  const int ni = bx.i1 - bx.i0;
  if (store == 0)
    hipLaunchKernelGGL(k_energy_planes<0>, dim3(ni), dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fields), sc,
                       mt, bx, part);
  else
    hipLaunchKernelGGL(k_energy_planes<1>, dim3(ni), dim3(256), 0, st, g, *static_cast<const Fields<1>*>(fields), sc,
                       mt, bx, part);
  note_launch("k_energy_planes");
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, part, ni, dv, out);
  note_launch("k_energy_final");
}

}  // namespace dtm
