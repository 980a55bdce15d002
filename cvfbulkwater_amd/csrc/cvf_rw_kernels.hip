// Histogram reweighting kernels: the per-cell log-sum over simulations of
// Calcula_Histograma (HistogramReweighting/histogram_reweighting_mixed_
// paramSI_norm_x.f90:1003-1124) for many targets in one launch, and the
// per-target max / log-sum-exp / normalization reductions.
//
// The reference's argument for cell c, target t and simulation j is
//   arg_j = beta_t (E + p_t V) - beta_j (E + p_j V) - c_j + lm_j = t1 + u_j,
// so the target only shifts a cell's sum: log sum_j exp(arg_j) =
// t1 + log sum_j exp(u_j).  k_rw_logp evaluates the n_sim terms u_j once
// per cell and then serves every target of its block with a few flops.
// The reference's two cut-offs are kept exactly:
//  - a cell's sum switches to the max-shifted form that drops terms more
//    than 10 below the largest once any |arg_j| > 10 (f90:1048-1078); the
//    switch depends on the target through t1 (t1 + max u > 10 or
//    t1 + min u < -10), the dropped set does not, so both sums (all terms,
//    and the terms within 10 of the largest) are formed once per cell;
//  - log_Z drops cells more than 10 below the largest log_P (1106-1118).
// Factoring t1 out rounds differently from the reference's left-to-right
// sum: log_P moves by about |t1| x 1e-16 (1e-10 for L = 8).
#include <math.h>

#include "cvf_rw_kernels.h"

namespace cvf {
namespace {

constexpr int kTargetsPerBlock = 16;  // targets served per k_rw_logp block
constexpr int kSumBlocks = 128;       // k_rw_sum blocks per target (grid-stride)

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// block-wide reduction of kRwBlock threads, result in every thread
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

// u_j = -beta_j (E + p_j V) - c_j + lm_j; s = {beta, p, c, lm}
__device__ inline double rw_u(double E, double V, const double* s) { return -(s[0] * (E + s[1] * V)) - s[2] + s[3]; }

__global__ __launch_bounds__(kRwBlock) void k_rw_logp(RwCells cells, const RwSim* __restrict__ sims,
                                                      const double* __restrict__ ctes, int32_t n_sim,
                                                      RwTargets tg, double* __restrict__ logp,
                                                      double* __restrict__ pmax, int32_t nblk) {
  extern __shared__ double sl[];  // 4 per simulation: beta, p, c, lm
  __shared__ double red[kTargetsPerBlock][kRwBlock / 64];
  for (int j = threadIdx.x; j < n_sim; j += kRwBlock) {
    sl[4 * j] = sims[j].beta;
    sl[4 * j + 1] = sims[j].pres;
    sl[4 * j + 2] = ctes[j];
    sl[4 * j + 3] = sims[j].lm;
  }
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * kRwBlock + threadIdx.x;
  const bool live = b < cells.n;
  double E = 0, V = 0, lh = 0, umax = -INFINITY, umin = INFINITY, all = 0, near = 0;
  if (live) {
    E = cells.en[b];
    V = cells.vn[b];
    lh = cells.lh[b];
    for (int j = 0; j < n_sim; ++j) {
      const double u = rw_u(E, V, sl + 4 * j);
      umax = fmax(umax, u);
      umin = fmin(umin, u);
    }
    double sa = 0, st = 0;
    for (int j = 0; j < n_sim; ++j) {
      const double d = rw_u(E, V, sl + 4 * j) - umax;
      const double e = exp(d);
      sa = sa + e;
      if (-d < 10.0) st = st + e;
    }
    all = log(sa) + umax;   // log sum_j exp(u_j)
    near = log(st) + umax;  // the same over the terms within 10 of the largest
  }
  // every target's wave maxima into LDS, then one barrier for all of them
  const int t0 = blockIdx.y * kTargetsPerBlock, nt = min(tg.n - t0, kTargetsPerBlock);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < nt; ++k) {
    const int t = t0 + k;
    double lp = -INFINITY;
    if (live) {
      const double2 bp = tg.bp[t];
      const double t1 = bp.x * (E + bp.y * V);
      const bool cut = t1 + umax > 10.0 || t1 + umin < -10.0;
      lp = lh - (t1 + (cut ? near : all));
      logp[(int64_t)t * cells.n + b] = lp;
    }
    lp = wave_max(lp);
    if (lane == 0) red[k][wave] = lp;
  }
  __syncthreads();
  if (threadIdx.x < nt) {
    double m = red[threadIdx.x][0];
    for (int w = 1; w < kRwBlock / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
    pmax[(int64_t)(t0 + threadIdx.x) * nblk + blockIdx.x] = m;
  }
}

// per target: the max of log_P from the k_rw_logp partials, then partial
// sums of exp(log_P - max) over the cells within 10 of the max (for log_Z)
// and over all cells (for the normalization), grid-stride over the cells
__global__ __launch_bounds__(kRwBlock) void k_rw_sum(const double* __restrict__ logp, int64_t n,
                                                     const double* __restrict__ pmax, int32_t nblk,
                                                     double* __restrict__ mx, double* __restrict__ ps) {
  __shared__ double red[kRwBlock / 64];
  const int t = blockIdx.y;
  const double* pm = pmax + (int64_t)t * nblk;
  double m = -INFINITY;
  for (int i = threadIdx.x; i < nblk; i += kRwBlock) m = fmax(m, pm[i]);
  m = block_reduce<true>(m, red);
  const double* lp = logp + (int64_t)t * n;
  double st = 0, sa = 0;
  for (int64_t b = (int64_t)blockIdx.x * kRwBlock + threadIdx.x; b < n; b += (int64_t)gridDim.x * kRwBlock) {
    const double v = lp[b];
    const double e = exp(v - m);
    sa = sa + e;
    if (m - v < 10.0) st = st + e;
  }
  st = block_reduce<false>(st, red);
  sa = block_reduce<false>(sa, red);
  if (threadIdx.x == 0) {
    ps[2 * ((int64_t)t * gridDim.x + blockIdx.x)] = st;
    ps[2 * ((int64_t)t * gridDim.x + blockIdx.x) + 1] = sa;
    if (blockIdx.x == 0) mx[t] = m;
  }
}

// one wave per target: log_Z = log(truncated sum) + max, the
// normalization sum exp(log_P - log_Z) de drho, and with ctes_out the
// constants' update log_Z(t) - log_Z(0)
__global__ __launch_bounds__(kRwBlock) void k_rw_finish(const double* __restrict__ ps, int32_t sblk, int32_t n_t,
                                                        const double* __restrict__ mx, double de, double drho,
                                                        double* __restrict__ lz, double* __restrict__ norm,
                                                        double* __restrict__ ctes_out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = wave; t < n_t; t += kRwBlock / 64) {
    double st = 0, sa = 0;
    for (int i = lane; i < sblk; i += 64) {
      st += ps[2 * ((int64_t)t * sblk + i)];
      sa += ps[2 * ((int64_t)t * sblk + i) + 1];
    }
    st = wave_sum(st);
    sa = wave_sum(sa);
    if (lane == 0) {
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

int32_t rw_sum_blocks(int32_t nblk) { return nblk < kSumBlocks ? nblk : kSumBlocks; }

// ------------------------------------------------------------ projections
// Each workgroup of a projection takes one contiguous range of cells, so
// every sum below runs in an order fixed by the launch shape alone.
__device__ inline void block_range(int64_t n, int64_t& lo, int64_t& hi) {
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  lo = (int64_t)blockIdx.x * per;
  hi = lo + per < n ? lo + per : n;
}

__device__ inline void cell_er(const RwGeom& g, uint32_t c, double& e, double& r) {
  const uint32_t i = c >> 16, w = c & 0xFFFFu;
  e = g.e_min + ((int)(i + 1) - 0.5) * g.de;
  r = g.rho_min + ((int)(w + 1) - 0.5) * g.drho;
}

constexpr int kRwBatch = 8;  // cells per thread whose loads a projection issues together

__global__ __launch_bounds__(kRwBlock) void k_rw_moments(RwGeom g, const double* __restrict__ dens,
                                                         double* __restrict__ part) {
  __shared__ double red[kRwBlock / 64];
  const int t = blockIdx.y;
  const double* v = dens + (int64_t)t * g.n;
  int64_t lo, hi;
  block_range(g.n, lo, hi);
  double s[5] = {0, 0, 0, 0, 0};
  // kRwBatch cells per thread per trip, their loads issued first, then
  // added in cell order
  for (int64_t b0 = lo + threadIdx.x; b0 < hi; b0 += kRwBlock * kRwBatch) {
    double xs[kRwBatch];
    uint32_t cs[kRwBatch];
#pragma unroll
    for (int u = 0; u < kRwBatch; ++u) {
      const int64_t bb = b0 + kRwBlock * u, bc = bb < hi ? bb : hi - 1;
      xs[u] = v[bc];
      cs[u] = g.cell[bc];
    }
#pragma unroll
    for (int u = 0; u < kRwBatch; ++u) {
      if (b0 + kRwBlock * u >= hi) break;
      double e, r;
      cell_er(g, cs[u], e, r);
      const double x = xs[u], e2 = e * e, r2 = r * r;
      s[0] += x;
      s[1] += e2 * x;
      s[2] += e2 * e2 * x;
      s[3] += r2 * x;
      s[4] += r2 * r2 * x;
    }
  }
  for (int k = 0; k < 5; ++k) {
    const double m = block_reduce<false>(s[k], red);
    if (threadIdx.x == 0) part[((int64_t)t * gridDim.x + blockIdx.x) * 5 + k] = m;
  }
}

__global__ __launch_bounds__(kRwBlock) void k_rw_xrange(RwGeom g, const double* __restrict__ dens, int32_t n_a,
                                                        const double2* __restrict__ sc,
                                                        double* __restrict__ part) {
  __shared__ double red[kRwBlock / 64];
  const int ta = blockIdx.y, t = ta / n_a, a = ta - t * n_a;
  const double* v = dens + (int64_t)t * g.n;
  const double2 q = sc[a];
  int64_t lo, hi;
  block_range(g.n, lo, hi);
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t b0 = lo + threadIdx.x; b0 < hi; b0 += kRwBlock * kRwBatch) {
    double xs[kRwBatch];
    uint32_t cs[kRwBatch];
#pragma unroll
    for (int u = 0; u < kRwBatch; ++u) {
      const int64_t bb = b0 + kRwBlock * u, bc = bb < hi ? bb : hi - 1;
      xs[u] = v[bc];
      cs[u] = g.cell[bc];
    }
#pragma unroll
    for (int u = 0; u < kRwBatch; ++u) {
      if (b0 + kRwBlock * u >= hi || !(xs[u] > 0)) continue;
      double e, r;
      cell_er(g, cs[u], e, r);
      const double x = e * q.x + r * q.y;
      mn = fmin(mn, x);
      mx = fmax(mx, x);
    }
  }
  mn = -block_reduce<true>(-mn, red);
  mx = block_reduce<true>(mx, red);
  if (threadIdx.x == 0) {
    part[2 * ((int64_t)ta * gridDim.x + blockIdx.x)] = mn;
    part[2 * ((int64_t)ta * gridDim.x + blockIdx.x) + 1] = mx;
  }
}

// One wavefront per workgroup; every lane adds its cells (lane, lane + 64,
// ... of the workgroup's contiguous range) into its own LDS row of nm bins,
// and lane k then sums column k over the 64 rows in lane order.  Every sum
// runs in an order fixed by the launch shape, with one LDS read-modify-write
// per cell and no cross-lane traffic.
constexpr int kRwXhistBatch = 16;

__global__ __launch_bounds__(64) void k_rw_xhist(RwGeom g, const double* __restrict__ dens, int32_t n_a,
                                                 const double2* __restrict__ sc,
                                                 const double2* __restrict__ binfo, double* __restrict__ part,
                                                 uint32_t* __restrict__ err) {
  extern __shared__ double hl[];  // [64 lanes][nm]
  const int ta = blockIdx.y, t = ta / n_a, a = ta - t * n_a;
  const int lane = threadIdx.x, nm = g.nm;
  double* h = hl + lane * nm;
  for (int k = 0; k < nm; ++k) h[k] = 0;
  const double* v = dens + (int64_t)t * g.n;
  const double2 q = sc[a], bi = binfo[ta];
  int64_t lo, hi;
  block_range(g.n, lo, hi);
  // kRwXhistBatch cells per lane per trip, all their loads issued first
  // (clamped indices), then added in cell order
  for (int64_t b0 = lo + lane; b0 < hi; b0 += 64 * kRwXhistBatch) {
    double xs[kRwXhistBatch];
    uint32_t cs[kRwXhistBatch];
#pragma unroll
    for (int u = 0; u < kRwXhistBatch; ++u) {
      const int64_t bb = b0 + 64 * u, bc = bb < hi ? bb : hi - 1;
      xs[u] = v[bc];
      cs[u] = g.cell[bc];
    }
#pragma unroll
    for (int u = 0; u < kRwXhistBatch; ++u) {
      const double x = xs[u];
      if (b0 + 64 * u >= hi || !(x > 0)) continue;
      double e, r;
      cell_er(g, cs[u], e, r);
      const double X = e * q.x + r * q.y, d = (X - bi.x) / bi.y;
      int k;
      if (fabs(d) < 1073741824.0 && (int)d >= 0 && (int)d < nm)
        k = (int)d;
      else if (fabs(X - bi.x) < (double)1E-10f)
        k = 0;
      else {
        atomicOr(err, 1u);
        continue;
      }
      h[k] += x;
    }
  }
  __syncthreads();
  for (int k = lane; k < nm; k += 64) {
    double s = 0;
    for (int l = 0; l < 64; ++l) s += hl[l * nm + k];
    part[((int64_t)ta * gridDim.x + blockIdx.x) * nm + k] = s;
  }
}

__global__ void k_rw_xhist_sum(const double* __restrict__ part, int32_t nm, double* __restrict__ out) {
  const int ta = blockIdx.x;
  for (int k = threadIdx.x; k < nm; k += blockDim.x) {
    double v = 0;
    for (int b = 0; b < kRwProjBlocks; ++b) v += part[((int64_t)ta * kRwProjBlocks + b) * nm + k];
    out[(int64_t)ta * nm + k] = v;
  }
}

}  // namespace

hipError_t launch_rw_xhist_sum(const double* part, int32_t n_ta, int32_t nm, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rw_xhist_sum, dim3((unsigned)n_ta), dim3(64), 0, st, part, nm, out);
  return hipGetLastError();
}

hipError_t launch_rw_moments(const RwGeom& g, const double* dens, int32_t n_t, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_rw_moments, dim3(kRwProjBlocks, (unsigned)n_t), dim3(kRwBlock), 0, st, g, dens, part);
  return hipGetLastError();
}

hipError_t launch_rw_xrange(const RwGeom& g, const double* dens, int32_t n_t, int32_t n_a, const double2* sc,
                            double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_rw_xrange, dim3(kRwProjBlocks, (unsigned)(n_t * n_a)), dim3(kRwBlock), 0, st, g, dens, n_a,
                     sc, part);
  return hipGetLastError();
}

hipError_t launch_rw_xhist(const RwGeom& g, const double* dens, int32_t n_t, int32_t n_a, const double2* sc,
                           const double2* binfo, double* part, uint32_t* err, hipStream_t st) {
  // 64 rows of nm <= 215 bins (n_bin2D <= 46340): up to 110 KB of LDS
  const size_t lds = sizeof(double) * 64 * (size_t)g.nm;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rw_xhist),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_rw_xhist, dim3(kRwProjBlocks, (unsigned)(n_t * n_a)), dim3(64), lds, st, g, dens, n_a, sc,
                     binfo, part, err);
  return hipGetLastError();
}

hipError_t launch_rw_pass(const RwCells& cells, const RwSim* sims, const double* ctes, int32_t n_sim,
                          const RwTargets& tg, const RwWork& w, double de, double drho, double* ctes_out,
                          bool dist, hipStream_t stream) {
  const unsigned rows = (unsigned)((tg.n + kTargetsPerBlock - 1) / kTargetsPerBlock);
  const size_t lds = sizeof(double) * 4 * (size_t)n_sim;
  hipLaunchKernelGGL(k_rw_logp, dim3((unsigned)w.nblk, rows), dim3(kRwBlock), lds, stream, cells, sims, ctes,
                     n_sim, tg, w.logp, w.pmax, w.nblk);
  const int32_t sblk = rw_sum_blocks(w.nblk);
  hipLaunchKernelGGL(k_rw_sum, dim3((unsigned)sblk, (unsigned)tg.n), dim3(kRwBlock), 0, stream, w.logp, cells.n,
                     w.pmax, w.nblk, w.mx, w.ps);
  hipLaunchKernelGGL(k_rw_finish, dim3(1), dim3(kRwBlock), 0, stream, w.ps, sblk, tg.n, w.mx, de, drho, w.lz,
                     w.norm, ctes_out);
  if (dist)
    hipLaunchKernelGGL(k_rw_dist, dim3((unsigned)w.nblk, (unsigned)tg.n), dim3(kRwBlock), 0, stream, w.logp,
                       cells.n, w.lz, w.norm);
  return hipGetLastError();
}

}  // namespace cvf
