// HIP kernels of the CVF bulk-water Monte Carlo engine (gfx950).
//
// The lattice state is one packed 32-bit word per cell (cvf_core.h).  One MC
// step of the reference schedule (MonteCarlo_Step, gpu_water3D.cu:1346-1733)
// is the launch sequence
//     step_begin -> eta(family 0) -> eta(family 1)
//                -> sigma(sub 0) ... sigma(sub 12*n_metropolis - 1)
//                -> [sw_local -> sw_union -> sw_flip] x n_cluster
// which the host captures once into a hipGraph and replays.  Every kernel
// takes a block table whose entry b carries the state point, the first work
// item and all per-point scalars of workgroup b, so one launch advances any
// number of state points of uneven L and a workgroup starts from a single
// scalar load.  Workgroup ids are remapped so that each XCD gets a
// contiguous range of the table (each XCD has its own L2).
#include "cvf_kernels.h"
#include "cvf_device.h"

#include <atomic>

#include <hip/hip_ext.h>

namespace cvf {

__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nb) {
  // hardware places workgroup b on XCD (b % 8); give XCD k the table range
  // [k*nb/8, (k+1)*nb/8).  nb is a multiple of 8.
  const uint32_t per = nb / kNumXcd;
  return (b % kNumXcd) * per + b / kNumXcd;
}

__device__ __forceinline__ uint64_t seed_of(const BlockEnt& e) {
  return ((uint64_t)e.seed_hi << 32) | e.seed_lo;
}

__device__ __forceinline__ uint32_t neighbor_up(uint32_t c, uint32_t coord, uint32_t stride,
                                                uint32_t L) {
  return coord == L - 1 ? c - (L - 1) * stride : c + stride;
}

// neighbour of cell c (coordinate `coord` along the arm's axis, cell stride
// `stride` along it) across arm `arm`
__device__ __forceinline__ uint32_t arm_neighbor(uint32_t c, uint32_t coord, uint32_t stride,
                                                 uint32_t L, int arm) {
  if (arm & 1) return coord == 0 ? c + (L - 1) * stride : c - stride;
  return coord == L - 1 ? c - (L - 1) * stride : c + stride;
}

__device__ __forceinline__ void cell_coords(const BlockEnt& e, uint32_t t, uint32_t& x, uint32_t& y,
                                            uint32_t& z) {
  const FastDiv dL{e.L, e.mL};
  const uint32_t r1 = fdiv(t, dL);
  x = t - r1 * e.L;
  z = fdiv(r1, dL);
  y = r1 - z * e.L;
}

// ------------------------------------------------------------ step begin
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float v) {
  // Neumaier's compensated step
  const float t = s + v;
  if (fabsf(s) >= fabsf(v))
    c += (s - t) + v;
  else
    c += (v - t) + s;
  s = t;
}

__device__ __forceinline__ void two_sum_pair(float& s, float& c, float s2, float c2) {
  const float t = s + s2;
  const float bp = t - s;
  const float err = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 + err;
}

// One Lennard-Jones term g = (r d)^-12 - (r d)^-6 - shift in FP32 as an
// unevaluated pair g_hi + g_lo, for r = r_hi + r_lo and shift = sh_hi +
// sh_lo, built with fma-based error-free transformations: x = r d, x^6,
// u = x^-6 and u (u - 1) - shift from their FP32 heads plus first-order
// FP32 corrections.
__device__ __forceinline__ void lj_term_fp32(float r_hi, float r_lo, float d, float sh_hi, float sh_lo, float& g_hi,
                                             float& g_lo) {
#pragma clang fp contract(off)
  // x = r d as x_hi + x_lo
  const float x_hi = r_hi * d;
  const float x_lo = fmaf(r_hi, d, -x_hi) + r_lo * d;
  // x^2, x^3, x^6 with first-order tails
  const float x2 = x_hi * x_hi;
  const float x2e = fmaf(x_hi, x_hi, -x2) + 2.f * x_hi * x_lo;
  const float x3 = x2 * x_hi;
  const float x3e = fmaf(x2, x_hi, -x3) + x2e * x_hi + x2 * x_lo;
  const float x6 = x3 * x3;
  const float x6e = fmaf(x3, x3, -x6) + 2.f * x3 * x3e;
  // u = 1/x^6 = u0 + u1
  const float u0 = 1.0f / x6;
  const float u1 = (fmaf(-x6, u0, 1.0f) - x6e * u0) * u0;
  // g = u (u - 1) - shift
  const float v0 = u0 - 1.0f;
  const float bp = v0 - u0;
  const float ev = (u0 - (v0 - bp)) + (-1.0f - bp);
  const float p0 = u0 * v0;
  const float p0e = fmaf(u0, v0, -p0);
  g_hi = p0 - sh_hi;
  const float gb = g_hi - p0;
  const float ge = (p0 - (g_hi - gb)) + (-sh_hi - gb);
  g_lo = ge + p0e + u0 * ev + u1 * (u0 + v0) - sh_lo;
}

// Per-molecule vdW sum with FP32 arithmetic.  Rounding r*d_k to a float
// alone would cost 6 x 6e-8 relative in (r d)^-6, so each term is carried as
// an unevaluated float pair (hi + lo) built with fma-based error-free
// transformations: x = r d_k, x^6, u = x^-6 and g = u (u - 1) - shift are
// formed from their FP32 heads plus first-order FP32 corrections.  The pairs
// are summed with Neumaier's compensated sum per lane and a compensated
// butterfly across the wavefront; the result is widened to FP64 once and
// scaled by N (the shell weight f_k / N = count_k / 2 is a small integer).
__device__ double vdw_fp32_wave(double r, const float* dist, const int32_t* freq, int nshell,
                                float r_cutoff, int N) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const double rc6 = 1.0 / ((double)r_cutoff * r_cutoff * r_cutoff * r_cutoff * r_cutoff * r_cutoff);
  const double shd = rc6 * rc6 - rc6;
  const float sh_hi = (float)shd, sh_lo = (float)(shd - (double)sh_hi);
  const float r_hi = (float)r, r_lo = (float)(r - (double)r_hi);
  float s = 0.f, c = 0.f;
  for (int k = lane; k < nshell; k += 64) {
    const float d = dist[k];
    if (r * (double)d < (double)r_cutoff) {
      float g_hi, g_lo;
      lj_term_fp32(r_hi, r_lo, d, sh_hi, sh_lo, g_hi, g_lo);
      // weight f_k / N = count_k / 2
      const float w = (float)(freq[k] / N);
      const float t_hi = w * g_hi;
      const float t_lo = fmaf(w, g_hi, -t_hi) + w * g_lo;
      two_sum_acc(s, c, t_hi);
      c += t_lo;
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const float s2 = __shfl_xor(s, m, 64);
    const float c2 = __shfl_xor(c, m, 64);
    two_sum_pair(s, c, s2, c2);
  }
  return ((double)s + (double)c) * (double)N;
}

// One wavefront per state point of a group: performs the point's volume move
// in device-volume mode (volume_step, gpu_water3D.cu:1740-1781), advances the
// point's step counter and publishes the step being executed for the
// group's kernels.  All points of a handle advance in lockstep.
// End of a captured sequence of nsteps MC steps of a group: the device
// volume moves of those steps (volume_step, gpu_water3D.cu:1740-1781), one
// wavefront per point running them in order from the point's own counter,
// and the group's step counter moved on.  The chain never reads the
// lattice, so running its moves after the lattice kernels of the same steps
// gives the same result as interleaving them.  No workgroup reads the group
// counter here, so workgroup 0 may advance it while others still run.
__global__ __launch_bounds__(64) void k_vol_chain(const PointDev* __restrict__ pts,
                                                  const int32_t* __restrict__ plist,
                                                  PointState* __restrict__ ps, uint64_t* __restrict__ gstep,
                                                  uint32_t nsteps, const float* __restrict__ dist,
                                                  const int32_t* __restrict__ freq, int precision,
                                                  int moves) {
  const int lane = threadIdx.x;
  if (blockIdx.x == 0 && lane == 0 && gstep) *gstep += nsteps;  // no counter: the moves alone
  if (!moves) return;
  const int p = plist[blockIdx.x];
  const PointDev& P = pts[p];
  const uint64_t step0 = ps[p].step;
  const float* d = dist + P.shell_off;
  const int32_t* f = freq + P.shell_off;
  VolState vs = ps[p].vol;
  for (uint32_t i = 0; i < nsteps; ++i) {
    double u2 = 0;
    const double trial = volume_trial(vs, P.vp, step0 + i, &u2);
    double te;
    if (precision == 1)
      te = vdw_fp32_wave(trial, d, f, P.vp.nshell, P.vp.r_cutoff, P.vp.N);
    else
      te = vdw_energy_fp64(trial, P.vp.r_cutoff, d, f, P.vp.nshell);
    // lane 0's energy decides, as one value for the whole wavefront, so every
    // lane keeps the same chain
    te = __shfl(te, 0);
    volume_accept(vs, P.vp, trial, te, u2);
  }
  if (lane == 0) {
    ps[p].vol = vs;
    ps[p].step = step0 + nsteps;
  }
}

// ------------------------------------------------------------ eta sweep
// Cube Metropolis of gpu_Chessboard_set_active_bonds (gpu_water3D.cu:
// 3374-3466).  Family 0 holds the cubes with all-even origin, family 1 the
// all-odd ones; within a family the cubes' corners are disjoint, so each
// thread updates its 8 corner words without atomics.  Cube states of both
// families are independent given sigma, so running the families one after
// the other equals the reference's simultaneous update.
// Four threads per quad of cubes of the family with origins
// (ox + b L/2, oy + a L/2, oz), a,b in {0,1}: the quad's Philox block (each
// of the four lanes forms it), word a + 2b for cube a,b, one cube per
// thread.  A workgroup covers blockDim/4 table items.  The corner loads are
// issued before the workgroup's barrier on the acceptance tables (lanes past
// the point's end load the entry's first quad and store nothing).
__global__ void k_eta(const BlockEnt* __restrict__ tab, uint32_t nb, const uint64_t* __restrict__ stepp,
                      uint32_t step_off,
                      const PointDev* __restrict__ pts, uint32_t* __restrict__ words, int family) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  const uint64_t step = *stepp + step_off;
  const uint32_t L = e.L, hL = L / 2, qL = L / 4;
  const uint32_t t = e.start + (threadIdx.x >> 2), q = threadIdx.x & 3u;
  const bool live = t < qL * qL * hL;
  const uint32_t tl = live ? t : e.start;
  const FastDiv dQ{qL, e.mQ};
  const uint32_t r1 = fdiv(tl, dQ);  // t = i + qL (j + qL k), i,j < L/4, k < L/2
  const uint32_t i = tl - r1 * qL;
  const uint32_t k = fdiv(r1, dQ);
  const uint32_t j = r1 - k * qL;
  const uint32_t ox = 2 * i + family, oy = 2 * j + family, oz = 2 * k + family;
  // cube q: (ox + (q & 2 ? L/2 : 0), oy + (q & 1 ? L/2 : 0), oz)
  const uint32_t cx = ox + ((q & 2u) ? hL : 0u), cy = oy + ((q & 1u) ? hL : 0u);
  const uint32_t x1 = cx + 1 == L ? 0 : cx + 1, y1 = cy + 1 == L ? 0 : cy + 1, z1 = oz + 1 == L ? 0 : oz + 1;
  uint32_t* W = words + e.woff;
  uint32_t idx[8], w[8];
#pragma unroll
  for (int b = 0; b < 8; ++b)
    idx[b] = ((b & 1) ? x1 : cx) + L * (((b & 2) ? y1 : cy) + L * ((b & 4) ? z1 : oz));
#pragma unroll
  for (int b = 0; b < 8; ++b) w[b] = W[idx[b]];
  __shared__ uint64_t thr[9];
  __shared__ uint64_t etab[9];
  if (threadIdx.x < 9) {
    thr[threadIdx.x] = pts[e.point].thr_eta[threadIdx.x];
    etab[threadIdx.x] = eta_table_entry(threadIdx.x);
  }
  __syncthreads();
  if (!live) return;
  const uint32_t cube_ref = ox / 2 + hL * (oy / 2) + hL * hL * oz;  // gpu_water3D.cu:3405
  const U4 r = draw(seed_of(e), kStEta, cube_ref, step, 0);
  const uint32_t rw = q == 0 ? r.x : q == 1 ? r.y : q == 2 ? r.z : r.w;
  eta_cube_move(w, thr, etab, rw, [&](int b, uint32_t v) { W[idx[b]] = v; });
}

// --------------------------------------------------------- sigma half-sweep
// gpu_update (gpu_water3D.cu:1890-2008), chessboard branch: arm `a` of the
// cells whose coordinate along a's axis is even (indices = 1) or odd
// (indices = 0).  A cell's new sigma_a depends on its own word and on the
// facing arm of the neighbour across arm a; that neighbour has the other
// parity along the axis, so no word read in a launch is written in it.
// The sigma rule for the four cells of a quad: words at W[cell[m]], facing
// neighbours at W[nb[m]] (LDS or global memory), random words from draw().
// The cells and their neighbours are distinct words (active cells against
// the other parity) and only the cells are written, so all eight loads are
// issued first, the Philox block is computed while they are in flight, and
// the four updates proceed without waiting on one another.
template <class Ptr, class Draw>
__device__ __forceinline__ void sigma_quad(Ptr W, const uint32_t (&cell)[4], const uint32_t (&nb)[4],
                                           const uint32_t* thr, int arm, Draw&& draw) {
  uint32_t w[4], wn[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    w[m] = W[cell[m]];
    wn[m] = W[nb[m]];
  }
  const U4 q = draw();
  uint32_t sp[4], ra[4], t[4];
  bool act[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t rw = m == 0 ? q.x : m == 1 ? q.y : m == 2 ? q.z : q.w;
    split_word(rw, 6, sp[m], ra[m]);
    const uint32_t sa = sigma_of(w[m], arm);
    act[m] = sp[m] != sa;
    t[m] = thr[sigma_thr_index(w[m], wn[m], arm, sa, sp[m])];
  }
#pragma unroll
  for (int m = 0; m < 4; ++m)
    if (act[m] && ra[m] < t[m]) W[cell[m]] = (w[m] & ~(7u << (3 * arm))) | (sp[m] << (3 * arm));
}

// One chessboard sigma move of arm `arm` on the word w held in a register
// (the rule of sigma_quad), facing neighbour word wn, random word rw.
__device__ __forceinline__ uint32_t sigma_move_reg(uint32_t w, uint32_t wn, const uint32_t* thr, int arm,
                                                   uint32_t rw) {
  uint32_t sp, ra;
  split_word(rw, 6, sp, ra);
  const uint32_t sa = sigma_of(w, arm);
  return sp != sa && ra < thr[sigma_thr_index(w, wn, arm, sa, sp)] ? (w & ~(7u << (3 * arm))) | (sp << (3 * arm))
                                                                    : w;
}

// Non-chessboard rule (gpu_water3D.cu:1966-1984, 2004-2007): with dN =
// delta(new, nb) - delta(old, nb), the HB term is
// [n_c + dN <= 4][n_nb + dN <= 4] delta(new, nb) - [n_c <= 4][n_nb <= 4]
// delta(old, nb), with the uint8 counters n in bits 24-31; an accepted move
// adds dN to both counters.  The neighbour has the other parity along the
// arm's axis, so no other thread of the launch reads or writes its word.
__device__ __forceinline__ void sigma_update_nc(uint32_t* __restrict__ W, const uint32_t* thr, uint32_t c,
                                                uint32_t nbc, int arm, uint32_t rw) {
  const uint32_t w = W[c];
  const uint32_t wn = W[nbc];
  const uint32_t sa = sigma_of(w, arm);
  uint32_t sp, ra;
  split_word(rw, 6, sp, ra);
  if (sp == sa) return;
  const uint32_t snb = sigma_of(wn, arm ^ 1);
  const int dn = (int)(sp == snb) - (int)(sa == snb);
  const int ni = (int)(w >> kNhbShift), nj = (int)(wn >> kNhbShift);
  const int h = ((ni + dn <= 4) && (nj + dn <= 4) && sp == snb) - ((ni <= 4) && (nj <= 4) && sa == snb);
  const int d = count_eq(w, sp) - count_eq(w, sa) + 1;
  if (ra < thr[(h + 1) * 11 + d + 5]) {
    W[c] = (w & kLowBits & ~(7u << (3 * arm))) | (sp << (3 * arm)) |
           ((uint32_t)((ni + dn) & 0xFF) << kNhbShift);
    if (dn) W[nbc] = (wn & kLowBits) | ((uint32_t)((nj + dn) & 0xFF) << kNhbShift);
  }
}

// The non-chessboard rule on the four cells of a quad (cells and facing
// neighbours all distinct words).
__device__ __forceinline__ void sigma_quad_nc(uint32_t* W, const uint32_t (&c)[4], const uint32_t (&nbc)[4],
                                              const uint32_t* thr, int arm, const U4& r) {
  sigma_update_nc(W, thr, c[0], nbc[0], arm, r.x);
  sigma_update_nc(W, thr, c[1], nbc[1], arm, r.y);
  sigma_update_nc(W, thr, c[2], nbc[2], arm, r.z);
  sigma_update_nc(W, thr, c[3], nbc[3], arm, r.w);
}

// One thread per quad of active cells (x + b L/2, y + a L/2, z), a,b in
// {0,1}, x,y < L/2: the four have the same parity along every axis (L/2 is
// even) and share one Philox block keyed on the a=b=0 cell (word a + 2b).
// The half-sweep's (arm, parity) is drawn in the kernel from uniform values
// (gpu_water3D.cu:1393-1394 draws them on the host).
__global__ void k_sigma(const BlockEnt* __restrict__ tab, uint32_t nb, const uint64_t* __restrict__ stepp,
                        uint32_t step_off,
                        const PointDev* __restrict__ pts, uint32_t* __restrict__ words, uint32_t sub) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  __shared__ uint32_t thr[33];
  if (threadIdx.x < 33) thr[threadIdx.x] = sigma_thr32(pts[e.point].thr_sigma[threadIdx.x]);
  __syncthreads();
  const uint32_t L = e.L, hL = L / 2, qL = L / 4, N = e.N;
  const uint32_t t = e.start + threadIdx.x;
  if (t >= N / 8) return;
  const uint64_t step = *stepp + step_off;
  const uint64_t seed = seed_of(e);
  const U4 sc = draw(seed, kStSched, 0, step, sub);
  const int arm = (int)uniform_int(sc.x, 6);
  const uint32_t q = (sc.y >> 31) ? 0u : 1u;  // kept parity along the axis
  const int ax = arm >> 1;
  const FastDiv dH{hL, e.mH}, dQ{qL, e.mQ};
  uint32_t x, y, z;  // quad origin, x,y < L/2
  if (ax == 0) {  // t = hx + qL (y + hL z)
    const uint32_t r1 = fdiv(t, dQ);
    x = 2 * (t - r1 * qL) + q;
    z = fdiv(r1, dH);
    y = r1 - z * hL;
  } else if (ax == 1) {  // t = x + hL (hy + qL z)
    const uint32_t r1 = fdiv(t, dH);
    x = t - r1 * hL;
    z = fdiv(r1, dQ);
    y = 2 * (r1 - z * qL) + q;
  } else {  // t = x + hL (y + hL hz)
    const uint32_t r1 = fdiv(t, dH);
    x = t - r1 * hL;
    const uint32_t hz = fdiv(r1, dH);
    y = r1 - hz * hL;
    z = 2 * hz + q;
  }
  const uint32_t dy = hL * L, dx = hL;
  const uint32_t c0 = x + L * (y + L * z);
  const uint32_t c[4] = {c0, c0 + dy, c0 + dx, c0 + dx + dy};
  uint32_t nbc[4];
  if (ax == 0) {
    nbc[0] = arm_neighbor(c[0], x, 1, L, arm);
    nbc[2] = arm_neighbor(c[2], x + hL, 1, L, arm);
    nbc[1] = nbc[0] + dy;
    nbc[3] = nbc[2] + dy;
  } else if (ax == 1) {
    nbc[0] = arm_neighbor(c[0], y, L, L, arm);
    nbc[1] = arm_neighbor(c[1], y + hL, L, L, arm);
    nbc[2] = nbc[0] + dx;
    nbc[3] = nbc[1] + dx;
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m) nbc[m] = arm_neighbor(c[m], z, L * L, L, arm);
  }
  uint32_t* W = words + e.woff;
  if (e.flags & kEntNoChess) {
    sigma_quad_nc(W, c, nbc, thr, arm, draw(seed, kStSigma, c0, step, sub));
    return;
  }
  sigma_quad(W, c, nbc, thr, arm, [&] { return draw(seed, kStSigma, c0, step, sub); });
}

// z-arm quad with origin (x, y) on the plane starting at `base`: the
// neighbour is on the plane above (arm 4) or below (arm 5)
__device__ __forceinline__ void quad_z_cells(uint32_t base, uint32_t L, int LL, uint32_t x, uint32_t y, int arm,
                                             uint32_t* cell, uint32_t* nb) {
  const uint32_t hL = L / 2;
  const int dz = (arm & 1) ? -LL : LL;
  const uint32_t b0 = base + x + L * y, dy = hL * L;
  cell[0] = b0;
  cell[1] = b0 + dy;
  cell[2] = b0 + hL;
  cell[3] = b0 + hL + dy;
  for (int m = 0; m < 4; ++m) nb[m] = cell[m] + dz;
}

// ------------------------------------------------------ fused slab kernel
// One workgroup per z-slab of T full x-y planes of one state point.  The
// slab plus a z-halo is staged in LDS, the optional eta sweep and nk
// (kChunk or kMaxChunk) consecutive sigma half-sweeps run there, and the slab's own planes are
// written to the other buffer of a ping-pong pair.  The halo is sized from
// the chunk's own schedule: walking the chunk backwards, every +z (-z)
// half-sweep widens the region that must be valid beforehand by one plane
// upward (downward), x/y half-sweeps stay in-plane (full periodic planes),
// and the eta sweep needs one more plane each side for its cubes.  Halo
// planes are recomputed redundantly; each workgroup computes exactly the
// updates the unfused kernels would make on those cells, with the same
// site-keyed random numbers, so the result is identical.
__device__ __forceinline__ uint32_t wrap_plane(int z, int L) {
  while (z < 0) z += L;
  while (z >= L) z -= L;
  return (uint32_t)z;
}

// Work items it = tid, tid + nt, ... enumerated as it = pi * Q + u (plane
// pi, in-plane item u) and advanced without a division per item.
struct PlaneIter {
  uint32_t pi, u;
};
__device__ __forceinline__ PlaneIter plane_first(uint32_t tid, uint32_t Q, FastDiv dq) {
  const uint32_t pi = fdiv(tid, dq);
  return PlaneIter{pi, tid - pi * Q};
}
__device__ __forceinline__ void plane_next(PlaneIter& it, uint32_t Q, uint32_t dp, uint32_t du) {
  it.u += du;
  it.pi += dp;
  if (it.u >= Q) {
    it.u -= Q;
    ++it.pi;
  }
}

constexpr int kMaxWin = kMaxSlab + 2 * kMaxChunk;  // window planes incl. halos (either rule)

// kAnyNc: the launch may hold non-chessboard slabs (compiled out otherwise).
template <bool kAnyNc>
__global__ void k_chunk(const BlockEnt* __restrict__ tab, uint32_t /* entries, launch order */,
                        const uint64_t* __restrict__ stepp,
                        uint32_t step_off,
                        const PointDev* __restrict__ pts, const uint32_t* __restrict__ src,
                        uint32_t* __restrict__ dst, uint32_t k0, uint32_t nk, int with_eta) {
  extern __shared__ __align__(16) uint32_t win[];
  __shared__ uint32_t thr_s[33];
  __shared__ uint64_t thr_e[9];
  __shared__ uint64_t etab[9];
  __shared__ uint32_t sch[kMaxChunk];
  __shared__ uint32_t gofs[kMaxWin];  // L^2 * global z of each window plane
  // the table is ordered by decreasing slab work and used in launch order,
  // so every XCD's dispatcher takes the large slabs first
  const BlockEnt e = tab[blockIdx.x];
  if (e.point < 0) return;
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const PointDev& P = pts[e.point];
  if (tid < 33) thr_s[tid] = sigma_thr32(P.thr_sigma[tid]);
  if (tid < 9) {
    thr_e[tid] = P.thr_eta[tid];
    etab[tid] = eta_table_entry(tid);
  }
  const uint64_t step = *stepp + step_off;
  const uint64_t seed = seed_of(e);
  if (tid < nk) {
    const U4 r = draw(seed, kStSched, 0, step, k0 + tid);
    sch[tid] = uniform_int(r.x, 6) | ((r.y >> 31) << 3);
  }
  __syncthreads();
  const int L = (int)e.L, hL = L / 2, qL = L / 4, LL = L * L, T = (int)e.T, z0 = (int)e.start;
  const FastDiv dH{(uint32_t)hL, e.mH}, dQ{(uint32_t)qL, e.mQ};
  const FastDiv dP8{(uint32_t)(LL / 8), e.mP8}, dP4{(uint32_t)(LL / 4), e.mP4},
      dC{(uint32_t)(qL * qL), e.mC};
  // planes relative to z0: window [wlo, whi).  A z half-sweep of the
  // non-chessboard rule also writes the HB counter of each updated cell's
  // facing neighbour, so it needs one more valid plane on both sides, and
  // such a point has no cube moves.
  const bool nc = kAnyNc && (e.flags & kEntNoChess) != 0;
  int h_lo = 0, h_hi = 0;
  for (uint32_t k = 0; k < nk; ++k) {
    const uint32_t arm = sch[k] & 7u;
    h_hi += nc ? arm >= 4 : arm == 4;
    h_lo += nc ? arm >= 4 : arm == 5;
  }
  const int ext = with_eta && !nc ? 1 : 0;
  const int wlo = -h_lo - ext, whi = T + h_hi + ext, np = whi - wlo;
  if (tid < (uint32_t)np) gofs[tid] = wrap_plane(z0 + wlo + (int)tid, L) * (uint32_t)LL;
  __syncthreads();
  // stage the window in 16-byte pieces (point offsets are 256-byte aligned
  // and L^2 is a multiple of 16 words), four loads in flight per thread
  {
    const uint4* S4 = reinterpret_cast<const uint4*>(src + e.woff);
    uint4* W4 = reinterpret_cast<uint4*>(win);
    const uint32_t Q = (uint32_t)(LL / 4), n = (uint32_t)np * Q;
    for (uint32_t i0 = tid; i0 < n; i0 += 4 * nt) {
      // pieces past the end redo the last one (same value, same place), so
      // the four loads are unconditional and issued together
      uint4 v[4];
      uint32_t ix[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ix[u] = min(i0 + (uint32_t)u * nt, n - 1);
        const uint32_t pi = mulhi32(ix[u], dP4.m);  // L^2/4 >= 4
        v[u] = S4[(gofs[pi] >> 2) + (ix[u] - pi * Q)];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) W4[ix[u]] = v[u];
    }
  }
  __syncthreads();
  if (ext) {
    // cubes with both planes inside the window, family by family
    const uint32_t Q = (uint32_t)(qL * qL), dp = fdiv(nt, dC), du = nt - dp * Q;
    for (int fam = 0; fam < 2; ++fam) {
      // first window plane whose global z has the family's parity
      const int first = ((z0 + wlo) & 1) == fam ? 0 : 1;
      const int nplanes = first < np - 1 ? (np - 1 - first + 1) / 2 : 0;
      for (PlaneIter it = plane_first(tid, Q, dC); it.pi < (uint32_t)nplanes; plane_next(it, Q, dp, du)) {
        const uint32_t jj = fdiv(it.u, dQ), ii = it.u - jj * qL;
        const int wz = first + 2 * (int)it.pi;
        const uint32_t ox = 2 * ii + fam, oy = 2 * jj + fam;
        // cube index x/2 + (L/2)(y/2) + (L^2/4) z (gpu_water3D.cu:3405)
        const uint32_t cube_ref = ii + hL * jj + (gofs[wz] >> 2);
        const U4 r = draw(seed, kStEta, cube_ref, step, 0);
        // cube origin plane wz; the cubes span wz, wz+1
        eta_cube_item(win, (uint32_t)wz * LL, (uint32_t)(wz + 1) * LL, thr_e, etab, (uint32_t)L, ox, oy, r);
      }
      __syncthreads();
    }
  }
  int rem_lo = h_lo, rem_hi = h_hi;
  for (uint32_t k = 0; k < nk; ++k) {
    const int arm = (int)(sch[k] & 7u);
    const uint32_t q = (sch[k] >> 3) ? 0u : 1u;
    const int ax = arm >> 1;
    rem_hi -= nc ? arm >= 4 : arm == 4;
    rem_lo -= nc ? arm >= 4 : arm == 5;
    // planes updated by this half-sweep; with the non-chessboard rule a +z
    // (-z) sweep also updates the active plane just below (above) the region
    // valid after it, for the counters it writes inside that region
    const int rlo = -rem_lo - (nc && arm == 4), rhi = T + rem_hi + (nc && arm == 5);
    const uint32_t sub = k0 + k;
    if (ax < 2) {
      const uint32_t Q = (uint32_t)(LL / 8), dp = fdiv(nt, dP8), du = nt - dp * Q;
      const uint32_t P = (uint32_t)(rhi - rlo), wz0 = (uint32_t)(rlo - wlo);
      for (PlaneIter it = plane_first(tid, Q, dP8); it.pi < P; plane_next(it, Q, dp, du)) {
        uint32_t x, y;
        if (ax == 0) {  // u = hx + qL y
          y = fdiv(it.u, dQ);
          x = 2 * (it.u - y * qL) + q;
        } else {  // u = x + hL hy
          const uint32_t hy = fdiv(it.u, dH);
          x = it.u - hy * hL;
          y = 2 * hy + q;
        }
        const uint32_t wz = wz0 + it.pi;
        uint32_t cell[4], nb[4];
        quad_plane_cells(wz * LL, (uint32_t)L, x, y, arm, ax, cell, nb);
        const auto rnd = [&] { return draw(seed, kStSigma, x + L * y + gofs[wz], step, sub); };
        if (nc)
          sigma_quad_nc(win, cell, nb, thr_s, arm, rnd());
        else
          sigma_quad(win, cell, nb, thr_s, arm, rnd);
      }
    } else {
      // planes whose global z has the kept parity
      const int first = rlo + ((((z0 + rlo) & 1) == (int)q) ? 0 : 1);
      const int nplanes = first < rhi ? (rhi - first + 1) / 2 : 0;
      const uint32_t Q = (uint32_t)(LL / 4), dp = fdiv(nt, dP4), du = nt - dp * Q;
      const uint32_t wz0 = (uint32_t)(first - wlo);
      for (PlaneIter it = plane_first(tid, Q, dP4); it.pi < (uint32_t)nplanes; plane_next(it, Q, dp, du)) {
        const uint32_t y = fdiv(it.u, dH), x = it.u - y * hL;
        const uint32_t wz = wz0 + 2 * it.pi;
        uint32_t cell[4], nb[4];
        quad_z_cells(wz * LL, (uint32_t)L, (int)LL, x, y, arm, cell, nb);
        const auto rnd = [&] { return draw(seed, kStSigma, x + L * y + gofs[wz], step, sub); };
        if (nc)
          sigma_quad_nc(win, cell, nb, thr_s, arm, rnd());
        else
          sigma_quad(win, cell, nb, thr_s, arm, rnd);
      }
    }
    __syncthreads();
  }
  // the slab's own planes go to the other buffer, 16-byte pieces
  uint4* D4 = reinterpret_cast<uint4*>(dst + e.woff + (uint32_t)z0 * LL);
  const uint4* Wi4 = reinterpret_cast<const uint4*>(win + (uint32_t)(-wlo) * LL);
  const uint32_t n4 = (uint32_t)(T * LL) / 4;
  for (uint32_t i0 = tid; i0 < n4; i0 += 4 * nt) {
    uint4 v[4];
    uint32_t ix[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ix[u] = min(i0 + (uint32_t)u * nt, n4 - 1);
      v[u] = Wi4[ix[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) D4[ix[u]] = v[u];
  }
}

// ------------------------------------------------------- fast schedule
// One colour pass of the fast schedule (CVF_SCHED_FAST, chessboard only):
// every cell of colour (x + y + z) % 2 updates its six arms in order, each
// with the single-move rule.  All neighbours of a cell have the other
// colour, so no word read by the pass is written by it; a thread keeps its
// quad's four words in registers through the six arms and writes them once.
// The quad (x + b L/2, y + a L/2, z), x, y < L/2, shares the colour (L/2 is
// even) and, per arm, one Philox block keyed on its a=b=0 cell (stream
// kStSigmaF, sub = 8 pass + arm; cell a,b takes word a + 2b).  Work item
// t = hx + (L/4) (y + (L/2) z) with x = 2 hx + ((y + z + colour) & 1):
// N/8 items per point, the half-sweep tables' count.
__global__ void k_sigma_f(const BlockEnt* __restrict__ tab, uint32_t nb, const uint64_t* __restrict__ stepp,
                          uint32_t step_off, const PointDev* __restrict__ pts, uint32_t* __restrict__ words,
                          uint32_t pass) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  __shared__ uint32_t thr[33];
  if (threadIdx.x < 33) thr[threadIdx.x] = sigma_thr32(pts[e.point].thr_sigma[threadIdx.x]);
  __syncthreads();
  const uint32_t L = e.L, hL = L / 2, qL = L / 4, N = e.N;
  const uint32_t t = e.start + threadIdx.x;
  if (t >= N / 8) return;
  const uint64_t step = *stepp + step_off;
  const uint64_t seed = seed_of(e);
  const FastDiv dH{hL, e.mH}, dQ{qL, e.mQ};
  const uint32_t r1 = fdiv(t, dQ), hx = t - r1 * qL;
  const uint32_t z = fdiv(r1, dH), y = r1 - z * hL;
  const uint32_t x = 2 * hx + ((y + z + (pass & 1u)) & 1u);
  const uint32_t dy = hL * L, dx = hL;
  const uint32_t c0 = x + L * (y + L * z);
  const uint32_t c[4] = {c0, c0 + dy, c0 + dx, c0 + dx + dy};
  const uint32_t cx[4] = {x, x, x + hL, x + hL}, cy[4] = {y, y + hL, y, y + hL};
  uint32_t* W = words + e.woff;
  uint32_t w[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) w[m] = W[c[m]];
  const uint32_t w0[4] = {w[0], w[1], w[2], w[3]};
#pragma unroll
  for (int arm = 0; arm < 6; ++arm) {
    const int ax = arm >> 1;
    uint32_t wn[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const uint32_t coord = ax == 0 ? cx[m] : ax == 1 ? cy[m] : z;
      const uint32_t stride = ax == 0 ? 1u : ax == 1 ? L : L * L;
      wn[m] = W[arm_neighbor(c[m], coord, stride, L, arm)];
    }
    const U4 r = draw(seed, kStSigmaF, c0, step, 8 * pass + (uint32_t)arm);
    w[0] = sigma_move_reg(w[0], wn[0], thr, arm, r.x);
    w[1] = sigma_move_reg(w[1], wn[1], thr, arm, r.y);
    w[2] = sigma_move_reg(w[2], wn[2], thr, arm, r.z);
    w[3] = sigma_move_reg(w[3], wn[3], thr, arm, r.w);
  }
#pragma unroll
  for (int m = 0; m < 4; ++m)
    if (w[m] != w0[m]) W[c[m]] = w[m];
}

// ----------------------------------------------------------- observables
// Integer part of measure() (gpu_water3D.cu:2014-2103), chessboard branch:
// 2*N_HB = sum over forward bonds with equal facing sigma of the two eta
// bits; N_sigma = sum over cells of equal-sigma arm pairs.
__global__ void k_zero_u32(uint32_t* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// three arrays zeroed by one launch (a, then b, then c in one index range)
__global__ void k_zero3_u32(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc) {
  const size_t n = na + nb + nc;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (i < na)
      a[i] = 0;
    else if (i < na + nb)
      b[i - na] = 0;
    else
      c[i - na - nb] = 0;
  }
}

// One workgroup of 256 threads per kObserveSpan cells of a point: each
// thread sums kObserveSpan / 256 cells, the workgroup reduces through LDS,
// and one pair of 64-bit atomics per workgroup adds to the point's totals
// (one pair per wavefront serialised on the same address at large L).
__global__ __launch_bounds__(256) void k_observe(const BlockEnt* __restrict__ tab, uint32_t nb,
                                                const uint32_t* __restrict__ words,
                                                unsigned long long* __restrict__ counters) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  const uint32_t L = e.L;
  const uint32_t* W = words + e.woff;
  uint32_t nhb2 = 0, nsig = 0;
  for (uint32_t k = 0; k < kObserveSpan / 256; ++k) {
    const uint32_t t = e.start + k * 256 + threadIdx.x;
    if (t >= e.N) break;
    uint32_t x, y, z;
    cell_coords(e, t, x, y, z);
    const uint32_t w = W[t];
    const uint32_t nbs[3] = {neighbor_up(t, x, 1, L), neighbor_up(t, y, L, L),
                             neighbor_up(t, z, L * L, L)};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const uint32_t wn = W[nbs[ax]];
      if (sigma_of(w, 2 * ax) == sigma_of(wn, 2 * ax + 1))
        nhb2 += eta_of(w, 2 * ax) + eta_of(wn, 2 * ax + 1);
    }
    uint32_t pairs = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) pairs += (uint32_t)count_eq(w, sigma_of(w, a)) - 1;
    nsig += pairs / 2;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    nhb2 += __shfl_xor(nhb2, m, 64);
    nsig += __shfl_xor(nsig, m, 64);
  }
  __shared__ uint32_t part[2][4];
  const uint32_t wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wv] = nhb2;
    part[1][wv] = nsig;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < 4; ++i) {
      a += part[0][i];
      b += part[1][i];
    }
    if (a | b) {
      atomicAdd(&counters[2 * e.point], a);
      atomicAdd(&counters[2 * e.point + 1], b);
    }
  }
}

// Counts of the six sigma values over the 6N arms of one point, as
// kHistReplicas partial histograms (counts[r * 6 + s], r = workgroup mod
// kHistReplicas) that the host sums: spreading the atomics keeps them from
// queueing on six addresses.  With a slot table (series, gstep, base, span)
// the counts go to the slot of the step just run, series + kHistWords
// (gstep + step_off - base), read on the device so that a captured step
// graph records every step it replays; slots outside [0, span) are not
// written.
__global__ __launch_bounds__(256) void k_sigma_hist(const uint32_t* __restrict__ W, uint32_t N,
                                                    unsigned long long* __restrict__ counts,
                                                    const uint64_t* __restrict__ gstep, uint32_t step_off,
                                                    const uint64_t* __restrict__ base, uint32_t span) {
  if (gstep) {
    const uint64_t slot = *gstep + step_off - *base;
    if (slot >= span) return;  // uniform over the grid
    counts += kHistWords * slot;
  }
  counts += 6 * (blockIdx.x % kHistReplicas);
  __shared__ uint32_t part[4][6];
  uint32_t c[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < N; t += gridDim.x * 256) {
    const uint32_t w = W[t];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const uint32_t v = sigma_of(w, a);
#pragma unroll
      for (int s = 0; s < 6; ++s) c[s] += v == (uint32_t)s;
    }
  }
#pragma unroll
  for (int s = 0; s < 6; ++s)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c[s] += __shfl_xor(c[s], m, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int s = 0; s < 6; ++s) part[threadIdx.x >> 6][s] = c[s];
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
    if (v) atomicAdd(&counts[threadIdx.x], (unsigned long long)v);
  }
}

// --------------------------------------------------------- Swendsen-Wang
// Bond activation of gpu_create_cluster (gpu_water3D.cu:2383-2438), then
// union-find with minimum-root linking.  Every link only ever points a
// larger node id to a smaller one, so each component's root is its minimum
// node id, the same canonical label the reference's iterative min-label
// propagation converges to (2440-2529), whatever the order of the atomics.
__device__ __forceinline__ uint32_t uf_load(const uint32_t* p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t uf_find(const uint32_t* par, uint32_t x) {
  uint32_t p = uf_load(par, x);
  while (p != x) {
    x = p;
    p = uf_load(par, x);
  }
  return x;
}

__device__ void uf_unite(uint32_t* par, uint32_t a, uint32_t b) {
  while (true) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t expected = a;
    if (__hip_atomic_compare_exchange_strong(par + a, &expected, b, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

// Links of one cell: 3 covalent (arm 0|2|4 of c with arm 1|3|5 of the +x|+y|+z
// neighbour) and 15 intra-molecule (arm pairs k<l in order).  Writes the
// intra-cell forest into `parent` and the active covalent links into cov[c].
__global__ void k_sw_local(const BlockEnt* __restrict__ tab, uint32_t nb,
                           const uint64_t* __restrict__ stepp, uint32_t step_off,
                           const uint32_t* __restrict__ words,
                           uint32_t* __restrict__ parent, uint8_t* __restrict__ cov, uint32_t jc) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  const uint32_t t = e.start + threadIdx.x;
  if (t >= e.N) return;
  const uint32_t L = e.L, N = e.N;
  const uint64_t step = *stepp + step_off;
  uint32_t x, y, z;
  cell_coords(e, t, x, y, z);
  const uint32_t* W = words + e.woff;
  const uint32_t w = W[t];
  const uint32_t nbs[3] = {neighbor_up(t, x, 1, L), neighbor_up(t, y, L, L),
                           neighbor_up(t, z, L * L, L)};
  uint32_t rw[20];
  const uint64_t seed = seed_of(e);
#pragma unroll
  for (int g = 0; g < 5; ++g) {
    const U4 r = draw(seed, kStSwBond, t, step, jc * 5 + g);
    rw[4 * g] = r.x;
    rw[4 * g + 1] = r.y;
    rw[4 * g + 2] = r.z;
    rw[4 * g + 3] = r.w;
  }
  const uint64_t thb = ((uint64_t)e.thb_hi << 32) | e.thb_lo;
  const uint64_t ts = ((uint64_t)e.ts_hi << 32) | e.ts_lo;
  uint32_t cmask = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const uint32_t wn = W[nbs[ax]];
    const uint32_t s0 = sigma_of(w, 2 * ax), s1 = sigma_of(wn, 2 * ax + 1);
    bool act = false;
    if (e.jeff_sign > 0)
      act = eta_of(w, 2 * ax) && s0 == s1 && (uint64_t)rw[ax] < thb;
    else if (e.jeff_sign < 0)
      act = s0 != s1 && (uint64_t)rw[ax] < thb;
    cmask |= (uint32_t)act << ax;
  }
  // intra-cell components of the 6 arms as bitmasks (Warshall closure in
  // registers); each arm's root is the lowest arm of its component
  uint32_t reach[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) reach[a] = 1u << a;
  int link = 3;
#pragma unroll
  for (int ii = 0; ii < 5; ++ii)
#pragma unroll
    for (int jj = ii + 1; jj < 6; ++jj) {
      if (sigma_of(w, ii) == sigma_of(w, jj) && (uint64_t)rw[link] < ts) {
        reach[ii] |= 1u << jj;
        reach[jj] |= 1u << ii;
      }
      ++link;
    }
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if ((reach[a] >> k) & 1u) reach[a] |= reach[k];
  uint32_t* par = parent + e.loff;
#pragma unroll
  for (int a = 0; a < 6; ++a) par[a * N + t] = (uint32_t)__builtin_ctz(reach[a]) * N + t;
  cov[e.woff + t] = (uint8_t)cmask;
}

__global__ void k_sw_union(const BlockEnt* __restrict__ tab, uint32_t nb, uint32_t* __restrict__ parent,
                           const uint8_t* __restrict__ cov) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  const uint32_t t = e.start + threadIdx.x;
  if (t >= e.N) return;
  const uint32_t cm = cov[e.woff + t];
  if (!cm) return;
  const uint32_t L = e.L, N = e.N;
  uint32_t x, y, z;
  cell_coords(e, t, x, y, z);
  const uint32_t nbs[3] = {neighbor_up(t, x, 1, L), neighbor_up(t, y, L, L),
                           neighbor_up(t, z, L * L, L)};
  uint32_t* par = parent + e.loff;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (cm & (1u << ax)) uf_unite(par, (2 * ax) * N + t, (2 * ax + 1) * N + nbs[ax]);
}

// Final labels (full path compression) and the cluster rotation of
// gpu_update_cluster (gpu_water3D.cu:2531-2542): sigma += floor(6u(label)).
__global__ void k_sw_flip(const BlockEnt* __restrict__ tab, uint32_t nb,
                          const uint64_t* __restrict__ stepp, uint32_t step_off,
                          uint32_t* __restrict__ words,
                          uint32_t* __restrict__ parent, uint32_t jc) {
  const BlockEnt e = tab[xcd_logical_block(blockIdx.x, nb)];
  if (e.point < 0) return;
  const uint32_t t = e.start + threadIdx.x;
  if (t >= e.N) return;
  const uint32_t N = e.N;
  const uint64_t step = *stepp + step_off;
  const uint64_t seed = seed_of(e);
  uint32_t* par = parent + e.loff;
  uint32_t* W = words + e.woff;
  uint32_t w = W[t];
  uint32_t nw = w & ~kSigmaBits;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    const uint32_t node = a * N + t;
    const uint32_t root = uf_find(par, node);
    __hip_atomic_store(par + node, root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const U4 r = draw(seed, kStSwFlip, root, step, jc);
    uint32_t s = sigma_of(w, a) + uniform_int(r.x, 6);
    s = s >= 6 ? s - 6 : s;
    nw |= s << (3 * a);
  }
  W[t] = nw;
}

// ------------------------------------------------- cluster statistics
// Nodes are visited in index order, 64 consecutive nodes per wavefront, so
// the nodes of one cluster come in runs of lanes (consecutive cells of one
// arm along x).  Each run is reduced inside the wavefront and its first lane
// makes one global atomic for it, instead of one atomic per node on the same
// address (a percolating cluster holds a large share of all 6N nodes).

// End of the run of lanes that starts at or before `lane`: the first lane
// above it that starts a run, or 64.  `heads` is the ballot of run starts.
__device__ __forceinline__ uint32_t run_end(uint64_t heads, uint32_t lane) {
  const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
  return above ? (uint32_t)__builtin_ctzll(above) : 64u;
}

__global__ __launch_bounds__(256) void k_cluster_sizes(const uint32_t* __restrict__ label, uint32_t nn,
                                                       int32_t* __restrict__ size) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63u;
  const uint32_t lab = i < nn ? label[i] : 0xFFFFFFFFu;  // no node has that label
  const uint32_t prev = __shfl_up(lab, 1, 64);
  const bool head = lane == 0 || lab != prev;
  const uint64_t heads = __ballot(head);
  if (head && i < nn) atomicAdd(&size[lab], (int32_t)(run_end(heads, lane) - lane));
}

// Order-preserving compaction of the cluster sizes (non-zero entries of
// size[], in label order): per workgroup of kCompactSpan labels the count
// of clusters, one workgroup's exclusive scan of those counts, then every
// workgroup writes its clusters at its offset.  The host then assembles the
// cluster.out row from the clusters alone, in the same order.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* sh, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    before += k < w ? sh[k] : 0u;
    all += sh[k];
  }
  *total = all;
  return before + x - v;
}

// Moments of the cluster sizes (count, sum, sum of squares, the largest two
// with repetition), reduced per workgroup and then over the workgroups.
// While every partial sum stays below 2^24 the reference's float
// accumulation of the sizes and of their squares in label order is exact
// (integers below 2^24 are floats), so these integers give its cluster.out
// row without the host loop.
struct Mom {
  uint64_t sum, sq;
  uint32_t m1, m2;  // m1 >= m2
};
__device__ __forceinline__ void mom_add(Mom& m, uint32_t v) {
  m.sum += v;
  m.sq += (uint64_t)v * v;
  if (v > m.m1) {
    m.m2 = m.m1;
    m.m1 = v;
  } else if (v > m.m2) {
    m.m2 = v;
  }
}
__device__ __forceinline__ void mom_merge(Mom& m, uint64_t sum, uint64_t sq, uint32_t a1, uint32_t a2) {
  m.sum += sum;
  m.sq += sq;
  const uint32_t n1 = max(m.m1, a1), n2 = max(min(m.m1, a1), max(m.m2, a2));
  m.m1 = n1;
  m.m2 = n2;
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int d) {
  const uint32_t lo = __shfl_xor((uint32_t)v, d, 64), hi = __shfl_xor((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}
// the workgroup's moments into out[0..3] (sum, sum of squares, m1 | m2 << 32)
// by thread 0; sh: 4 x 4 words of LDS
__device__ __forceinline__ void mom_block_store(Mom m, uint64_t* sh, uint64_t* out) {
  for (int d = 32; d >= 1; d >>= 1) {
    const uint64_t s2 = shfl_xor_u64(m.sum, d), q2 = shfl_xor_u64(m.sq, d);
    const uint32_t a1 = __shfl_xor(m.m1, d, 64), a2 = __shfl_xor(m.m2, d, 64);
    mom_merge(m, s2, q2, a1, a2);
  }
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[4 * w] = m.sum;
    sh[4 * w + 1] = m.sq;
    sh[4 * w + 2] = m.m1 | ((uint64_t)m.m2 << 32);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Mom t{0, 0, 0, 0};
    for (uint32_t k = 0; k < blockDim.x / 64; ++k)
      mom_merge(t, sh[4 * k], sh[4 * k + 1], (uint32_t)sh[4 * k + 2], (uint32_t)(sh[4 * k + 2] >> 32));
    out[0] = t.sum;
    out[1] = t.sq;
    out[2] = t.m1 | ((uint64_t)t.m2 << 32);
    out[3] = 0;
  }
}

__global__ __launch_bounds__(256) void k_compact_count(const int32_t* __restrict__ size, uint32_t nn,
                                                       uint32_t* __restrict__ bcount, uint64_t* __restrict__ bmom) {
  __shared__ uint32_t sh[4];
  __shared__ uint64_t shm[16];
  const uint32_t base = blockIdx.x * kCompactSpan + threadIdx.x * 4;
  uint32_t c = 0;
  Mom m{0, 0, 0, 0};
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t v = base + k < nn ? (uint32_t)size[base + k] : 0u;
    c += v != 0;
    mom_add(m, v);
  }
  uint32_t total;
  block_excl_scan256(c, sh, &total);
  if (threadIdx.x == 0) bcount[blockIdx.x] = total;
  mom_block_store(m, shm, bmom + 4 * blockIdx.x);
}

// one workgroup: bcount[0..nb) -> exclusive offsets, the sum into *total;
// the workgroups' moments bmom[0..nb) into bmom[nb]
__global__ __launch_bounds__(256) void k_compact_scan(uint32_t* __restrict__ bcount, uint32_t nb,
                                                      uint32_t* __restrict__ total, uint64_t* __restrict__ bmom) {
  __shared__ uint32_t sh[4];
  __shared__ uint64_t shm[16];
  uint32_t run = 0;
  Mom m{0, 0, 0, 0};
  for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < nb ? bcount[i] : 0u;
    if (i < nb) mom_merge(m, bmom[4 * i], bmom[4 * i + 1], (uint32_t)bmom[4 * i + 2], (uint32_t)(bmom[4 * i + 2] >> 32));
    uint32_t sum;
    const uint32_t ex = block_excl_scan256(v, sh, &sum);
    if (i < nb) bcount[i] = run + ex;
    run += sum;
    __syncthreads();  // sh is reused by the next round
  }
  if (threadIdx.x == 0) *total = run;
  mom_block_store(m, shm, bmom + 4 * nb);
}

__global__ __launch_bounds__(256) void k_compact_write(const int32_t* __restrict__ size, uint32_t nn,
                                                       const uint32_t* __restrict__ boff, int32_t* __restrict__ out) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * kCompactSpan + threadIdx.x * 4;
  int32_t v[4];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    v[k] = base + k < nn ? size[base + k] : 0;
    c += v[k] != 0;
  }
  uint32_t total;
  uint32_t pos = boff[blockIdx.x] + block_excl_scan256(c, sh, &total);
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k)
    if (v[k]) out[pos++] = v[k];
}

__global__ void k_perc_slots(const uint32_t* __restrict__ label, uint32_t nn,
                             const int32_t* __restrict__ size, int32_t minsize,
                             int32_t* __restrict__ slot, uint32_t* __restrict__ nslot,
                             uint32_t max_slots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn && label[i] == i && size[i] >= minsize) {
    const uint32_t s = atomicAdd(nslot, 1u);
    slot[i] = s < max_slots ? (int32_t)s : -1;
  }
}

// Coverage bitsets of the percolation test (is_percolating_cluster,
// gpu_water3D.cu:2723-2780): for each cluster of >= minsize nodes (its slot),
// the x, y and z values its cells take.  Lanes of one run share the slot and
// the bitset words their coordinates fall in; the run ORs its bits together
// (segmented butterfly over the wavefront) and its first lane makes three
// atomics.
__global__ __launch_bounds__(256) void k_perc_cover(const uint32_t* __restrict__ label, uint32_t nn, uint32_t L,
                                                    const int32_t* __restrict__ size, int32_t minsize,
                                                    const int32_t* __restrict__ slot, uint32_t* __restrict__ cover,
                                                    uint32_t max_slots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63u;
  uint32_t key = 0xFFFFFFFFu, bx = 0, by = 0, bz = 0, x = 0, y = 0, z = 0;
  if (i < nn) {
    const uint32_t r = label[i];
    const int32_t s = size[r] >= minsize ? slot[r] : -1;
    if (s >= 0 && (uint32_t)s < max_slots) {
      const uint32_t N = L * L * L;
      const uint32_t cell = i % N;
      x = cell % L;
      y = (cell / L) % L;
      z = cell / (L * L);
      // slot < 3 L^2 + 1 < 2^18, word indices < 8 (L <= 252)
      key = ((uint32_t)s << 9) | ((x >> 5) << 6) | ((y >> 5) << 3) | (z >> 5);
      bx = 1u << (x & 31u);
      by = 1u << (y & 31u);
      bz = 1u << (z & 31u);
    }
  }
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || key != prev;
  const uint32_t end = run_end(__ballot(head), lane);
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t ox = __shfl_down(bx, d, 64), oy = __shfl_down(by, d, 64), oz = __shfl_down(bz, d, 64);
    if (lane + d < end) {
      bx |= ox;
      by |= oy;
      bz |= oz;
    }
  }
  if (!head || key == 0xFFFFFFFFu) return;
  const uint32_t nw = (L + 31) / 32;
  uint32_t* cv = cover + (size_t)(key >> 9) * 3 * nw;
  atomicOr(&cv[x >> 5], bx);
  atomicOr(&cv[nw + (y >> 5)], by);
  atomicOr(&cv[2 * nw + (z >> 5)], bz);
}

__global__ void k_perc_count(const uint32_t* __restrict__ cover, const uint32_t* __restrict__ nslot,
                             uint32_t L, uint32_t max_slots, uint32_t* __restrict__ nperc) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ns = *nslot < max_slots ? *nslot : max_slots;
  if (s >= ns) return;
  const uint32_t nw = (L + 31) / 32;
  const uint32_t* cv = cover + (size_t)s * 3 * nw;
  bool any = false;
  for (int ax = 0; ax < 3; ++ax) {
    bool full = true;
    for (uint32_t k = 0; k < nw; ++k) {
      const uint32_t bits = (k == nw - 1 && (L % 32)) ? ((1u << (L % 32)) - 1u) : 0xFFFFFFFFu;
      if ((cv[ax * nw + k] & bits) != bits) full = false;
    }
    any = any || full;
  }
  if (any) atomicAdd(nperc, 1u);
}

// ------------------------------------------------------ vdW evaluation
__global__ __launch_bounds__(64) void k_vdw_eval(const float* __restrict__ dist,
                                                 const int32_t* __restrict__ freq, int nshell,
                                                 float r_cutoff, int N, const double* __restrict__ r,
                                                 int precision, double* __restrict__ out) {
  const double rr = r[blockIdx.x];
  double e;
  if (precision == 1)
    e = vdw_fp32_wave(rr, dist, freq, nshell, r_cutoff, N);
  else
    e = vdw_energy_fp64(rr, r_cutoff, dist, freq, nshell);
  if (threadIdx.x == 0) out[blockIdx.x] = e;
}

// ------------------------------------------------ vdW pair-form validation
// The pair form of calculate_energy_vdW (gpu_water3D.cu:1784-1805): every
// molecule i against every other molecule j whose minimum-image distance
// (distance_PBC, 1323-1337: float of the root of the folded integer
// displacement) is below R_cutoff (the shell table's filter, 886) and
// scaled by r below R_cutoff (the energy's filter).  No shell table is
// used: the sum runs over pairs from the lattice coordinates.  One thread
// per molecule i over the offsets of the window [-R, R]^3 (R = floor
// R_cutoff, clipped to the lattice's periodic range so that every j
// appears once); each workgroup writes its partial sum (FP64, or the FP32
// compensated pair hi, lo) for the radius of grid row blockIdx.y.
__device__ __forceinline__ int fold_pbc(int a, int L) {
  a = a < 0 ? -a : a;
  return a > L / 2 ? L - a : a;
}

__global__ __launch_bounds__(256) void k_vdw_pairs(int L, int R, float r_cutoff, const double* __restrict__ rr,
                                                   int precision, double* __restrict__ part) {
#pragma clang fp contract(off)
  const int N = L * L * L;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double r = rr[blockIdx.y];
  const double rc6 = 1.0 / ((double)r_cutoff * r_cutoff * r_cutoff * r_cutoff * r_cutoff * r_cutoff);
  double tmp = 1.0 / (double)r_cutoff;
  tmp = tmp * tmp * tmp;
  const double att0 = tmp * tmp;
  const double shift = att0 * att0 - att0;
  const double shd = rc6 * rc6 - rc6;
  const float sh_hi = (float)shd, sh_lo = (float)(shd - (double)sh_hi);
  const float r_hi = (float)r, r_lo = (float)(r - (double)r_hi);
  double sd = 0;
  float s = 0.f, c = 0.f;
  if (i < N) {
    const int x1 = i % L, y1 = (i / L) % L, z1 = i / (L * L);
    // offsets cover each j once: (-L/2, L/2] along an axis of the lattice
    const int lo = -min(R, L / 2 - 1), hi = min(R, L / 2);
    for (int oz = lo; oz <= hi; ++oz)
      for (int oy = lo; oy <= hi; ++oy)
        for (int ox = lo; ox <= hi; ++ox) {
          if (!(ox | oy | oz)) continue;
          const int x2 = (x1 + ox + L) % L, y2 = (y1 + oy + L) % L, z2 = (z1 + oz + L) % L;
          const int dx = fold_pbc(x1 - x2, L), dy = fold_pbc(y1 - y2, L), dz = fold_pbc(z1 - z2, L);
          const float d = (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
          if (!(d < r_cutoff) || !(r * (double)d < (double)r_cutoff)) continue;
          if (precision == 1) {
            float g_hi, g_lo;
            lj_term_fp32(r_hi, r_lo, d, sh_hi, sh_lo, g_hi, g_lo);
            two_sum_acc(s, c, g_hi);
            c += g_lo;
          } else {
            double t = 1.0 / (r * d);
            t = t * t * t;
            const double att = t * t;
            sd += att * att - att - shift;
          }
        }
  }
  __shared__ double ps[4];
  __shared__ float pf[4][2];
  if (precision == 1) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const float s2 = __shfl_xor(s, m, 64), c2 = __shfl_xor(c, m, 64);
      two_sum_pair(s, c, s2, c2);
    }
    if ((threadIdx.x & 63) == 0) {
      pf[threadIdx.x >> 6][0] = s;
      pf[threadIdx.x >> 6][1] = c;
    }
  } else {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) sd += __shfl_xor(sd, m, 64);
    if ((threadIdx.x & 63) == 0) ps[threadIdx.x >> 6] = sd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < 4; ++w) t += precision == 1 ? (double)pf[w][0] + (double)pf[w][1] : ps[w];
    part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
  }
}

// ------------------------------------------------------------ launchers
static inline dim3 grid1(size_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

// When set, the next launches record their own execution time into the two
// events (hipExtLaunchKernelGGL timestamps the dispatch itself).
static thread_local hipEvent_t g_t0 = nullptr, g_t1 = nullptr;
void set_launch_timing(hipEvent_t start, hipEvent_t stop) {
  g_t0 = start;
  g_t1 = stop;
}
void get_launch_timing(hipEvent_t* start, hipEvent_t* stop) {
  *start = g_t0;
  *stop = g_t1;
}

template <typename K, typename... A>
static void launch(K kernel, dim3 grid, dim3 block, hipStream_t st, A... args) {
  if (g_t0)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, g_t0, g_t1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

void launch_vol_chain(hipStream_t st, const PointDev* pts, const int32_t* plist, int npts, PointState* ps,
                      uint64_t* gstep, uint32_t nsteps, const float* dist, const int32_t* freq,
                      int precision, int moves) {
  launch(k_vol_chain, dim3(moves ? npts : 1), dim3(64), st, pts, plist, ps, gstep, nsteps, dist, freq,
         precision, moves);
}
void launch_eta(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, int family) {
  launch(k_eta, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, ln.step, ln.step_off, pts, words,
         family);
}
void launch_sigma(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, uint32_t sub) {
  launch(k_sigma, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, ln.step, ln.step_off, pts, words,
         sub);
}
void launch_sigma_f(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, uint32_t pass) {
  launch(k_sigma_f, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, ln.step, ln.step_off, pts, words,
         pass);
}
uint32_t chunk_lds_limit() {
  // gfx950: 160 KB of LDS per CU; keep room for the static arrays
  return 160 * 1024 - 1024;
}
template <bool kAnyNc>
void launch_chunk_t(hipStream_t st, Launch ln, const PointDev* pts, const uint32_t* src, uint32_t* dst, uint32_t k0,
                    uint32_t nk, int with_eta, uint32_t lds_bytes) {
  // the dynamic-LDS limit is a per-device function attribute: set it once
  // on each device the library launches on
  static std::atomic<uint64_t> attr_set{0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t bit = 1ull << (dev & 63);
  if (!(attr_set.load(std::memory_order_acquire) & bit)) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_chunk<kAnyNc>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)chunk_lds_limit());
    attr_set.fetch_or(bit, std::memory_order_acq_rel);
  }
  if (g_t0)
    hipExtLaunchKernelGGL(k_chunk<kAnyNc>, dim3(ln.nblocks), dim3(ln.block), lds_bytes, st, g_t0, g_t1, 0,
                          ln.tab, ln.nblocks, ln.step, ln.step_off, pts, src, dst, k0, nk, with_eta);
  else
    hipLaunchKernelGGL(k_chunk<kAnyNc>, dim3(ln.nblocks), dim3(ln.block), lds_bytes, st, ln.tab, ln.nblocks,
                       ln.step, ln.step_off, pts, src, dst, k0, nk, with_eta);
}
void launch_chunk(hipStream_t st, Launch ln, const PointDev* pts, const uint32_t* src, uint32_t* dst,
                  uint32_t k0, uint32_t nk, int with_eta, uint32_t lds_bytes, bool any_nc) {
  if (any_nc)
    launch_chunk_t<true>(st, ln, pts, src, dst, k0, nk, with_eta, lds_bytes);
  else
    launch_chunk_t<false>(st, ln, pts, src, dst, k0, nk, with_eta, lds_bytes);
}
void launch_zero_u32(hipStream_t st, uint32_t* p, size_t n) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  launch(k_zero_u32, dim3((unsigned)blocks), dim3(256), st, p, n);
}
void launch_zero3_u32(hipStream_t st, uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc) {
  size_t blocks = (na + nb + nc + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  launch(k_zero3_u32, dim3((unsigned)blocks), dim3(256), st, a, na, b, nb, c, nc);
}
void launch_observe(hipStream_t st, Launch ln, const uint32_t* words, unsigned long long* counters) {
  launch(k_observe, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, words, counters);
}
void launch_sw_local(hipStream_t st, Launch ln, const uint32_t* words, uint32_t* parent, uint8_t* cov,
                     uint32_t jc) {
  launch(k_sw_local, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, ln.step, ln.step_off, words,
         parent, cov, jc);
}
void launch_sw_union(hipStream_t st, Launch ln, uint32_t* parent, const uint8_t* cov) {
  launch(k_sw_union, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, parent, cov);
}
void launch_sw_flip(hipStream_t st, Launch ln, uint32_t* words, uint32_t* parent, uint32_t jc) {
  launch(k_sw_flip, dim3(ln.nblocks), dim3(ln.block), st, ln.tab, ln.nblocks, ln.step, ln.step_off, words,
         parent, jc);
}
// workgroups of the sigma histogram: about 4 cells per thread, at most 256
static dim3 hist_grid(uint32_t N) { return dim3(std::max(1u, std::min(256u, (N + 1023) / 1024))); }
void launch_sigma_hist(hipStream_t st, const uint32_t* W, uint32_t N, unsigned long long* counts) {
  launch(k_sigma_hist, hist_grid(N), dim3(256), st, W, N, counts, (const uint64_t*)nullptr, 0u,
         (const uint64_t*)nullptr, 0u);
}
void launch_sigma_hist_slot(hipStream_t st, const uint32_t* W, uint32_t N, unsigned long long* series,
                            const uint64_t* gstep, uint32_t step_off, const uint64_t* base, uint32_t span) {
  launch(k_sigma_hist, hist_grid(N), dim3(256), st, W, N, series, gstep, step_off, base, span);
}
void launch_cluster_sizes(hipStream_t st, const uint32_t* label, uint32_t nn, int32_t* size) {
  launch(k_cluster_sizes, grid1(nn, 256), dim3(256), st, label, nn, size);
}
uint32_t compact_blocks(uint32_t nn) { return (nn + kCompactSpan - 1) / kCompactSpan; }
void launch_compact_count(hipStream_t st, const int32_t* size, uint32_t nn, uint32_t* bcount, uint32_t* total,
                          uint64_t* bmom) {
  const uint32_t nb = compact_blocks(nn);
  launch(k_compact_count, dim3(nb), dim3(256), st, size, nn, bcount, bmom);
  launch(k_compact_scan, dim3(1), dim3(256), st, bcount, nb, total, bmom);
}
void launch_compact_write(hipStream_t st, const int32_t* size, uint32_t nn, const uint32_t* bcount, int32_t* out) {
  launch(k_compact_write, dim3(compact_blocks(nn)), dim3(256), st, size, nn, bcount, out);
}
void launch_perc_slots(hipStream_t st, const uint32_t* label, uint32_t nn, const int32_t* size,
                       int32_t minsize, int32_t* slot, uint32_t* nslot, uint32_t max_slots) {
  launch(k_perc_slots, grid1(nn, 256), dim3(256), st, label, nn, size, minsize, slot, nslot, max_slots);
}
void launch_perc_cover(hipStream_t st, const uint32_t* label, uint32_t nn, uint32_t L,
                       const int32_t* size, int32_t minsize, const int32_t* slot, uint32_t* cover,
                       uint32_t max_slots) {
  launch(k_perc_cover, grid1(nn, 256), dim3(256), st, label, nn, L, size, minsize, slot, cover,
         max_slots);
}
void launch_perc_count(hipStream_t st, const uint32_t* cover, const uint32_t* nslot, uint32_t L,
                       uint32_t max_slots, uint32_t* nperc) {
  launch(k_perc_count, grid1(max_slots, 256), dim3(256), st, cover, nslot, L, max_slots, nperc);
}
void launch_vdw_eval(hipStream_t st, const float* dist, const int32_t* freq, int nshell,
                     float r_cutoff, int N, const double* r, int n_r, int precision, double* out) {
  launch(k_vdw_eval, dim3(n_r), dim3(64), st, dist, freq, nshell, r_cutoff, N, r, precision, out);
}
uint32_t vdw_pair_blocks(int L) { return (uint32_t)(((size_t)L * L * L + 255) / 256); }
void launch_vdw_pairs(hipStream_t st, int L, float r_cutoff, const double* r, int n_r, int precision, double* part) {
  launch(k_vdw_pairs, dim3(vdw_pair_blocks(L), n_r), dim3(256), st, L, (int)r_cutoff, r_cutoff, r, precision, part);
}

}  // namespace cvf
