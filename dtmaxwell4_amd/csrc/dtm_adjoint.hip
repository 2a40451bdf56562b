// dtm_adjoint.hip — device side of the design-gradient adjoint.
//
// Objective: J = sum_m |F_m|^2, F_m = sum_{n=0}^{N-1} Ez_m^{n+1} e^{i w (n+1) dt} dt
// over monitor cells m (the running DFT of the probes), differentiated with
// respect to the inverse permittivity of the design box.
//
// The leapfrog step (H^{n+1/2} = H^{n-1/2} - dt Ce E^n, E^{n+1} = E^n + dt M Ch
// H^{n+1/2}, M = 1/eps, Ch = Ce^T for the (2,4) staggered differences) has
// the transposed step
//   lH~ = lH + dt Ce (M lE^{n+1}),   lE^n = lE^{n+1} - dt Ch lH~ + dJ/dE^n,
// which in zeta = M lE and eta = -lH~ is the forward step itself:
//   eta -= dt Ce zeta,   zeta += dt M Ch eta + M dJ/dE^n.
// So the adjoint runs the forward sweeps backward in n on (zeta, eta), with
// the objective's derivative injected at the monitors, and
//   dJ/dM_c = sum_n lE_c^{n+1} dt (Ch H^{n+1/2})_c = sum_n eps_c^2 zeta_c^{n+1} (E_c^{n+1} - E_c^n)
// from the design-box E history of the forward run.  (The CPML is applied in
// its forward form to the adjoint fields, an absorbing layer for them too;
// the reference has no adjoint.)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "dtm_internal.h"

namespace dtm {

namespace {

// E of the design box, three components, into out[c * nd + cell].
__global__ __launch_bounds__(256) void k_adj_snap(Grid g, Fields<0> f, Materials mt, float* out) {
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  const long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (b >= nd) return;
  const int k = mt.d_k0 + (int)(b % mt.d_nk);
  const int j = mt.d_j0 + (int)((b / mt.d_nk) % mt.d_nj);
  const int i = mt.d_i0 + (int)(b / ((long long)mt.d_nk * mt.d_nj));
  const long long q = i * g.sx + j * g.sy + k;
  for (int c = 0; c < 3; c++) out[c * nd + b] = f.E[c][q];
}

// series[m] = Ez at the monitor cells.
__global__ __launch_bounds__(256) void k_adj_monitor(Fields<0> f, const long long* q, int nmon, float* series) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nmon) series[m] = f.E[2][q[m]];
}

// acc[cell] += sum_c eps^2 zeta_c (E1_c - E0_c)  (design inverse permittivity is isotropic)
__global__ __launch_bounds__(256) void k_adj_grad(Grid g, Fields<0> f, Materials mt, const float* e0, const float* e1,
                                                  double* acc) {
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  const long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (b >= nd) return;
  const int k = mt.d_k0 + (int)(b % mt.d_nk);
  const int j = mt.d_j0 + (int)((b / mt.d_nk) % mt.d_nj);
  const int i = mt.d_i0 + (int)(b / ((long long)mt.d_nk * mt.d_nj));
  const long long q = i * g.sx + j * g.sy + k;
  const double eps = 1.0 / (double)mt.d_inv_eps[b];
  double s = 0.0;
  for (int c = 0; c < 3; c++) s += (double)f.E[c][q] * ((double)e1[c * nd + b] - (double)e0[c * nd + b]);
  acc[b] += eps * eps * s;
}

// zeta_z += M dJ/dEz^n at the monitors: dJ/dEz_m^n = 2 dt Re(conj(F_m) e^{i w n dt}).
__global__ __launch_bounds__(256) void k_adj_inject(Fields<0> f, const long long* q, const float* inv_eps,
                                                    const double* F, int nmon, double w, double dt, long long n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmon) return;
  double sn, cn;
  sincos(w * (double)n * dt, &sn, &cn);
  const double gJ = 2.0 * dt * (F[2 * m] * cn + F[2 * m + 1] * sn);
  f.E[2][q[m]] += (float)(gJ * inv_eps[m]);
}

// k_adj_inject at step n and k_adj_grad of step n-1 in one launch: the
// first nbg blocks accumulate the gradient, the rest inject.  Only used when
// no monitor lies in the design box, so the two touch disjoint cells.
__global__ __launch_bounds__(256) void k_adj_inject_grad(Grid g, Fields<0> f, Materials mt, const float* e0,
                                                         const float* e1, double* acc, unsigned nbg,
                                                         const long long* q, const float* inv_eps, const double* F,
                                                         int nmon, double w, double dt, long long n) {
  if (blockIdx.x < nbg) {
    const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
    const long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (b >= nd) return;
    const int k = mt.d_k0 + (int)(b % mt.d_nk);
    const int j = mt.d_j0 + (int)((b / mt.d_nk) % mt.d_nj);
    const int i = mt.d_i0 + (int)(b / ((long long)mt.d_nk * mt.d_nj));
    const long long qq = i * g.sx + j * g.sy + k;
    const double eps = 1.0 / (double)mt.d_inv_eps[b];
    double sum = 0.0;
    for (int c = 0; c < 3; c++) sum += (double)f.E[c][qq] * ((double)e1[c * nd + b] - (double)e0[c * nd + b]);
    acc[b] += eps * eps * sum;
    return;
  }
  const int m = (blockIdx.x - nbg) * blockDim.x + threadIdx.x;
  if (m >= nmon) return;
  double sn, cn;
  sincos(w * (double)n * dt, &sn, &cn);
  const double gJ = 2.0 * dt * (F[2 * m] * cn + F[2 * m + 1] * sn);
  f.E[2][q[m]] += (float)(gJ * inv_eps[m]);
}

void check(const char* what) { note_launch(what); }

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_adj_snap(const Grid& g, const void* fields, const Materials& mt, float* out, hipStream_t st) {
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  hipLaunchKernelGGL(k_adj_snap, dim3(blocks_for(nd)), dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fields), mt,
                     out);
  check("k_adj_snap");
}

void launch_adj_monitor(const void* fields, const long long* q, int nmon, float* series, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_monitor, dim3(blocks_for(nmon)), dim3(256), 0, st, *static_cast<const Fields<0>*>(fields),
                     q, nmon, series);
  check("k_adj_monitor");
}

void launch_adj_grad(const Grid& g, const void* fields, const Materials& mt, const float* e0, const float* e1,
                     double* acc, hipStream_t st) {
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  hipLaunchKernelGGL(k_adj_grad, dim3(blocks_for(nd)), dim3(256), 0, st, g, *static_cast<const Fields<0>*>(fields), mt,
                     e0, e1, acc);
  check("k_adj_grad");
}

void launch_adj_inject_grad(const Grid& g, const void* fields, const Materials& mt, const float* e0, const float* e1,
                            double* acc, const long long* q, const float* inv_eps, const double* F, int nmon,
                            double w, double dt, long long n, hipStream_t st) {
  const long long nd = (long long)mt.d_ni * mt.d_nj * mt.d_nk;
  const unsigned nbg = blocks_for(nd);
  hipLaunchKernelGGL(k_adj_inject_grad, dim3(nbg + blocks_for(nmon)), dim3(256), 0, st, g,
                     *static_cast<const Fields<0>*>(fields), mt, e0, e1, acc, nbg, q, inv_eps, F, nmon, w, dt, n);
  check("k_adj_inject_grad");
}

void launch_adj_inject(const void* fields, const long long* q, const float* inv_eps, const double* F, int nmon,
                       double w, double dt, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_inject, dim3(blocks_for(nmon)), dim3(256), 0, st, *static_cast<const Fields<0>*>(fields), q,
                     inv_eps, F, nmon, w, dt, n);
  check("k_adj_inject");
}

}  // namespace dtm
