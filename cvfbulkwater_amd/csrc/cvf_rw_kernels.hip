// Histogram reweighting kernels: the per-cell log-sum over simulations of
// Calcula_Histograma (HistogramReweighting/histogram_reweighting_mixed_
// paramSI_norm_x.f90:1003-1124) for many targets in one launch, and the
// per-target max / log-sum-exp / normalization reductions.
//
// Every cell and target is independent, so one thread takes one (cell,
// target) pair; the simulations' parameters sit in LDS and every lane of a
// wave reads the same entry (a broadcast).  The arithmetic is FP64 in the
// reference's order, including its two cut-offs: a cell's sum over
// simulations switches to the max-shifted form with terms below max - 10
// dropped once any |arg| exceeds 10, and log_Z drops cells more than 10
// below the maximum.
#include <math.h>

#include "cvf_rw_kernels.h"

namespace cvf {
namespace {

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// block-wide reduction of kRwBlock threads; result valid in thread 0
template <bool kMax>
__device__ inline double block_reduce(double v, double* red) {
  v = kMax ? wave_max(v) : wave_sum(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < kRwBlock / 64; ++i) r = kMax ? fmax(r, red[i]) : r + red[i];
  __syncthreads();
  return r;
}

// arg_j of f90:1048: beta1 (E + p1 V) - beta_j (E + p_j V) - c_j + lm_j
__device__ inline double rw_arg(double t1, double E, double V, const double* s) {
  return t1 - s[0] * (E + s[1] * V) - s[2] + s[3];
}

__global__ __launch_bounds__(kRwBlock) void k_rw_logp(RwCells cells, const RwSim* __restrict__ sims,
                                                      const double* __restrict__ ctes, int32_t n_sim,
                                                      RwTargets tg, double* __restrict__ logp,
                                                      double* __restrict__ pmax, int32_t nblk) {
  extern __shared__ double sl[];  // 4 per simulation: beta, p, c, lm
  __shared__ double red[kRwBlock / 64];
  for (int j = threadIdx.x; j < n_sim; j += kRwBlock) {
    sl[4 * j] = sims[j].beta;
    sl[4 * j + 1] = sims[j].pres;
    sl[4 * j + 2] = ctes[j];
    sl[4 * j + 3] = sims[j].lm;
  }
  __syncthreads();
  const int t = blockIdx.y;
  const int64_t b = (int64_t)blockIdx.x * kRwBlock + threadIdx.x;
  double lp = -INFINITY;
  if (b < cells.n) {
    const double E = cells.en[b], V = cells.vn[b];
    const double2 bp = tg.bp[t];
    const double t1 = bp.x * (E + bp.y * V);
    bool flag = false;
    double mx = rw_arg(t1, E, V, sl);
    for (int j = 0; j < n_sim; ++j) {
      const double a = rw_arg(t1, E, V, sl + 4 * j);
      flag |= fabs(a) > 10.0;
      if (mx < a) mx = a;
    }
    double xx = 0;
    if (!flag) {
      for (int j = 0; j < n_sim; ++j) xx = xx + exp(rw_arg(t1, E, V, sl + 4 * j));
      xx = log(xx);
    } else {
      for (int j = 0; j < n_sim; ++j) {
        const double a = rw_arg(t1, E, V, sl + 4 * j);
        if (mx - a < 10.0) xx = xx + exp(a - mx);
      }
      if (xx != 0) xx = log(xx);
      xx = xx + mx;
    }
    lp = cells.lh[b] - xx;
    logp[(int64_t)t * cells.n + b] = lp;
  }
  const double m = block_reduce<true>(lp, red);
  if (threadIdx.x == 0) pmax[(int64_t)t * nblk + blockIdx.x] = m;
}

__global__ __launch_bounds__(kRwBlock) void k_rw_max(const double* __restrict__ pmax, int32_t nblk,
                                                     double* __restrict__ mx) {
  __shared__ double red[kRwBlock / 64];
  const double* p = pmax + (int64_t)blockIdx.x * nblk;
  double m = -INFINITY;
  for (int i = threadIdx.x; i < nblk; i += kRwBlock) m = fmax(m, p[i]);
  m = block_reduce<true>(m, red);
  if (threadIdx.x == 0) mx[blockIdx.x] = m;
}

// partial sums of exp(log_P - max): over cells within 10 of the max (for
// log_Z) and over all cells (for the normalization)
__global__ __launch_bounds__(kRwBlock) void k_rw_sum(const double* __restrict__ logp, int64_t n,
                                                     const double* __restrict__ mx, int32_t nblk,
                                                     double* __restrict__ ps) {
  __shared__ double red[kRwBlock / 64];
  const int t = blockIdx.y;
  const int64_t b = (int64_t)blockIdx.x * kRwBlock + threadIdx.x;
  double st = 0, sa = 0;
  if (b < n) {
    const double m = mx[t], v = logp[(int64_t)t * n + b];
    const double e = exp(v - m);
    sa = e;
    if (m - v < 10.0) st = e;
  }
  st = block_reduce<false>(st, red);
  sa = block_reduce<false>(sa, red);
  if (threadIdx.x == 0) {
    ps[2 * ((int64_t)t * nblk + blockIdx.x)] = st;
    ps[2 * ((int64_t)t * nblk + blockIdx.x) + 1] = sa;
  }
}

__global__ __launch_bounds__(kRwBlock) void k_rw_finish(const double* __restrict__ ps, int32_t nblk, int32_t n_t,
                                                        const double* __restrict__ mx, double de, double drho,
                                                        double* __restrict__ lz, double* __restrict__ norm,
                                                        double* __restrict__ ctes_out) {
  __shared__ double red[kRwBlock / 64];
  for (int t = 0; t < n_t; ++t) {
    double st = 0, sa = 0;
    for (int i = threadIdx.x; i < nblk; i += kRwBlock) {
      st += ps[2 * ((int64_t)t * nblk + i)];
      sa += ps[2 * ((int64_t)t * nblk + i) + 1];
    }
    st = block_reduce<false>(st, red);
    sa = block_reduce<false>(sa, red);
    if (threadIdx.x == 0) {
      const double z = (st != 0 ? log(st) : 0.0) + mx[t];
      lz[t] = z;
      norm[t] = sa * exp(mx[t] - z) * de * drho;
    }
  }
  __syncthreads();
  if (ctes_out)
    for (int t = threadIdx.x; t < n_t; t += kRwBlock) ctes_out[t] = lz[t] - lz[0];
}

__global__ __launch_bounds__(kRwBlock) void k_rw_dist(double* __restrict__ logp, int64_t n,
                                                      const double* __restrict__ lz,
                                                      const double* __restrict__ norm) {
  const int t = blockIdx.y;
  const int64_t b = (int64_t)blockIdx.x * kRwBlock + threadIdx.x;
  if (b < n) {
    double* p = logp + (int64_t)t * n + b;
    *p = exp(*p - lz[t]) / norm[t];
  }
}

}  // namespace

hipError_t launch_rw_pass(const RwCells& cells, const RwSim* sims, const double* ctes, int32_t n_sim,
                          const RwTargets& tg, const RwWork& w, double de, double drho, double* ctes_out,
                          bool dist, hipStream_t stream) {
  const dim3 grid((unsigned)w.nblk, (unsigned)tg.n);
  const size_t lds = sizeof(double) * 4 * (size_t)n_sim;
  hipLaunchKernelGGL(k_rw_logp, grid, dim3(kRwBlock), lds, stream, cells, sims, ctes, n_sim, tg, w.logp, w.pmax,
                     w.nblk);
  hipLaunchKernelGGL(k_rw_max, dim3((unsigned)tg.n), dim3(kRwBlock), 0, stream, w.pmax, w.nblk, w.mx);
  hipLaunchKernelGGL(k_rw_sum, grid, dim3(kRwBlock), 0, stream, w.logp, cells.n, w.mx, w.nblk, w.ps);
  hipLaunchKernelGGL(k_rw_finish, dim3(1), dim3(kRwBlock), 0, stream, w.ps, w.nblk, tg.n, w.mx, de, drho, w.lz,
                     w.norm, ctes_out);
  if (dist) hipLaunchKernelGGL(k_rw_dist, grid, dim3(kRwBlock), 0, stream, w.logp, cells.n, w.lz, w.norm);
  return hipGetLastError();
}

}  // namespace cvf
