// Resident Metropolis kernel: the lattice stays in LDS for a whole sequence
// of MC steps (the reference schedule: eta cube sweep + 12 sigma
// half-sweeps per step, gpu_water3D.cu:1346-1417, 1890-2008, 3374-3466).
//
// Each state point is cut into ns z-slabs of T planes (ns = 1: the whole
// lattice in one window).  The host deals the slabs to at most one
// workgroup per CU, balancing cells, so every workgroup of the launch is
// resident at once.  A workgroup stages its slabs once, runs `nsteps` MC
// steps on them in LDS and writes them back once.  A slab's window holds
// its T planes plus a ghost copy of the plane below (window plane 0) and of
// the plane above (window plane T + 1), owned by the neighbouring slabs.
//
// Only z half-sweeps and the cube sweep read across a slab boundary:
//   * a +z half-sweep (arm 4) of the top plane reads sigma_5 of the plane
//     above, which changes only in -z half-sweeps of that plane, and a -z
//     half-sweep of the bottom plane reads sigma_4 of the plane below;
//   * a cube whose origin plane is the last plane of a slab (or the ghost
//     below) spans the boundary; both slabs compute it, from the same words
//     and the same cube-keyed random numbers, so both copies agree.
// A slab therefore publishes its bottom (top) plane after every -z (+z)
// half-sweep in which that plane is active, and both boundary planes at the
// end of every step, into rings in global memory: slot (step & 1) * 16 + j
// for the j-th publish of the step, slot (step & 1) * 16 + 15 for the end
// of the step, each followed by a version ((step + 1) << 4 | j).  A reader
// waits for the version it needs, then reads the slot.
//
// Coherence across the XCDs' L2 caches: every ring word and every version
// is accessed with agent-scope relaxed atomics (single-location coherent
// loads and stores that bypass the XCDs' non-coherent L2 caching), so no
// L2 writeback or invalidation is needed.  Order: a writer's ring stores
// complete (s_waitcnt, then the workgroup barrier) before its version
// store is issued; a reader issues its ring loads only after its version
// load has returned a large enough value.
// A neighbour can be at most one step ahead (each step starts with both
// neighbours' end-of-step planes), so the two halves of the ring are never
// overwritten before they are read.
//
// Every wait is bounded: a wait that times out, or sees another workgroup's
// abort, sets the abort flag in `status`; from then on every workgroup skips
// its remaining work and waits (its barriers still run, uniformly), writes
// nothing back, and the host reports the error.
#include "cvf_device.h"
#include "cvf_kernels.h"

#include <atomic>

#include <hip/hip_ext.h>

namespace cvf {

namespace {

constexpr uint32_t kSpinLimit = 1u << 24;  // polls of one wait (~1 s with s_sleep 2)

__device__ __forceinline__ uint64_t ring_version(uint64_t step, uint32_t j) { return ((step + 1) << 4) | j; }

__device__ __forceinline__ uint32_t ring_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ring_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until *v >= want.  False on time-out or abort.
__device__ bool wait_version(const uint64_t* v, uint64_t want, uint32_t* status) {
  if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
  for (uint32_t it = 0;; ++it) {
    if (__hip_atomic_load(v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) {
      // no later load moves above the one that saw the version
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return true;
    }
    if ((it & 63u) == 63u && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if (it >= kSpinLimit) {
      __hip_atomic_fetch_or(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Store a version after the workgroup barrier that follows the slot's ring
// stores (each writing wave waited for its stores to complete before it).
__device__ __forceinline__ void publish(uint64_t* v, uint64_t version) {
  __hip_atomic_store(v, version, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A writing wave's ring stores are complete (vmcnt 0) before it reaches the
// barrier after which the version is published.
__device__ __forceinline__ void ring_stores_done() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
}

__device__ __forceinline__ uint32_t parity_of(int z, uint32_t L) { return (uint32_t)(z + (int)L) & 1u; }

// Cube move on the cube with origin (ox, oy) of bottom plane B and top
// plane U (L x L words each, periodic in x and y).
__device__ __forceinline__ void eta_cube_planes(uint32_t* B, uint32_t* U, const uint64_t* thr, const uint64_t* etab,
                                                uint32_t L, uint32_t ox, uint32_t oy, uint32_t rw) {
  const uint32_t x1 = ox + 1 == L ? 0 : ox + 1;
  const uint32_t y1 = oy + 1 == L ? 0 : oy + 1;
  uint32_t* P[8];
  uint32_t w[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t x = (b & 1) ? x1 : ox, y = (b & 2) ? y1 : oy;
    P[b] = ((b & 4) ? U : B) + x + L * y;
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) w[b] = *P[b];
  eta_cube_move(w, thr, etab, rw, [&](int b, uint32_t v) { *P[b] = v; });
}

// Chessboard sigma rule on four words held in registers (facing neighbour
// words wn, one Philox block), as sigma_quad.
__device__ __forceinline__ void sigma_quad_regs(uint32_t (&w)[4], const uint32_t (&wn)[4], const uint32_t* thr,
                                                int arm, const U4& q) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t rw = m == 0 ? q.x : m == 1 ? q.y : m == 2 ? q.z : q.w;
    uint32_t sp, ra;
    split_word(rw, 6, sp, ra);
    const uint32_t sa = sigma_of(w[m], arm);
    const uint32_t t = thr[sigma_thr_index(w[m], wn[m], arm, sa, sp)];
    const uint32_t nw = (w[m] & ~(7u << (3 * arm))) | (sp << (3 * arm));
    w[m] = sp != sa && ra < t ? nw : w[m];
  }
}

// Items of one slab in one phase of a step: planes first, first + step,
// ... (np of them), Q items per plane.
struct Seg {
  int32_t first;  // first slab-relative plane
  uint32_t np;    // planes
  uint32_t Q;     // items per plane
  uint32_t mQ;    // fast-division multiplier of Q
  uint32_t step;  // plane stride (1 or 2)
};

constexpr int kPhases = 14;  // per step: two cube families, then the 12 half-sweeps

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-uniform copy of the fields of the slab a wave's items belong to.
struct Cur {
  uint32_t g, L, hL, qL, LL, z0, T, ns, lds, mH, mQ;
  uint64_t seed;
  Seg sg;
};

// The slab of 64-item chunk ch of phase ph (pre: the slabs' first chunks)
// and its fields, as scalars: every chunk lies within one slab.
__device__ __forceinline__ Cur chunk_slab(const uint32_t* pre, const Seg* seg, const RSlab* S, uint32_t ch) {
  Cur c;
  uint32_t g = 0;
  while (ch >= pre[g + 1]) ++g;
  c.g = uni(g);
  const RSlab& s = S[c.g];
  c.L = uni(s.L);
  c.hL = c.L / 2;
  c.qL = c.L / 4;
  c.LL = c.L * c.L;
  c.z0 = uni(s.z0);
  c.T = uni(s.T);
  c.ns = uni(s.ns);
  c.lds = uni(s.lds);
  c.mH = uni(s.mH);
  c.mQ = uni(s.mQ);
  c.seed = ((uint64_t)uni(s.seed_hi) << 32) | uni(s.seed_lo);
  const Seg& sg = seg[c.g];
  c.sg = Seg{(int32_t)uni((uint32_t)sg.first), uni(sg.np), uni(sg.Q), uni(sg.mQ), uni(sg.step)};
  return c;
}

}  // namespace

// grid = one workgroup per RGroupEnt; dynamic LDS = the largest sum of
// window sizes of one workgroup.
__global__ __launch_bounds__(1024) void k_resident(const RGroupEnt* __restrict__ grp, const RSlab* __restrict__ slabs,
                                                    const PointDev* __restrict__ pts, uint32_t* __restrict__ words,
                                                    uint32_t* __restrict__ rings, uint64_t* __restrict__ ver,
                                                    const uint64_t* __restrict__ stepp, uint32_t step_off,
                                                    uint32_t nsteps, uint32_t* __restrict__ status) {
  extern __shared__ __align__(16) uint32_t lds[];
  __shared__ RSlab S[kMaxRSlabs];
  __shared__ uint32_t thr_s[kMaxRSlabs][33];
  __shared__ uint64_t thr_e[kMaxRSlabs][9];
  __shared__ uint64_t etab[9];
  __shared__ uint32_t sch[kMaxRSlabs][12];
  // per slab and half-sweep: ring slot j + 1 to read above (+z, top plane)
  // or below (-z, bottom plane), 0 = the window's ghost; slot j + 1 to
  // publish into, 0 = none
  __shared__ uint8_t rd[kMaxRSlabs][12], wr[kMaxRSlabs][12];
  // items of each slab in each phase of the step, and their prefix sums
  __shared__ Seg seg[kPhases][kMaxRSlabs];
  __shared__ uint32_t pre[kPhases][kMaxRSlabs + 1];
  __shared__ uint32_t abort_s;
  const RGroupEnt G = grp[blockIdx.x];
  const uint32_t nsl = G.count, tid = threadIdx.x, nt = blockDim.x;
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(slabs + G.first);
    uint32_t* dst = reinterpret_cast<uint32_t*>(S);
    for (uint32_t i = tid; i < nsl * (uint32_t)(sizeof(RSlab) / 4); i += nt) dst[i] = src[i];
  }
  if (tid == 0) abort_s = 0;
  if (tid < 9) etab[tid] = eta_table_entry(tid);
  __syncthreads();
  for (uint32_t i = tid; i < nsl * 33; i += nt) {
    const uint32_t g = i / 33, k = i - g * 33;
    thr_s[g][k] = sigma_thr32(pts[S[g].point].thr_sigma[k]);
  }
  for (uint32_t i = tid; i < nsl * 9; i += nt) {
    const uint32_t g = i / 9, k = i - g * 9;
    thr_e[g][k] = pts[S[g].point].thr_eta[k];
  }
  // stage every window in 16-byte pieces (L^2 is a multiple of 16 words)
  for (uint32_t g = 0; g < nsl; ++g) {
    const RSlab& s = S[g];
    const uint32_t LL = s.L * s.L, Q4 = LL / 4, np = s.ns > 1 ? s.T + 2 : s.T;
    const uint4* W4 = reinterpret_cast<const uint4*>(words + s.woff);
    uint4* D4 = reinterpret_cast<uint4*>(lds + s.lds);
    for (uint32_t i = tid; i < np * Q4; i += nt) {
      const uint32_t p = i / Q4, r = i - p * Q4;
      const int zg = s.ns > 1 ? (int)s.z0 - 1 + (int)p : (int)p;
      const uint32_t z = (uint32_t)(zg < 0 ? zg + (int)s.L : zg >= (int)s.L ? zg - (int)s.L : zg);
      D4[i] = W4[z * Q4 + r];
    }
  }
  __syncthreads();
  const uint64_t step0 = *stepp + step_off;
  for (uint32_t si = 0; si < nsteps; ++si) {
    const uint64_t step = step0 + si;
    const uint32_t half = (uint32_t)(step & 1u) * kRingHalf;
    if (tid < nsl * 12) {
      const uint32_t g = tid / 12, k = tid - g * 12;
      const uint64_t seed = ((uint64_t)S[g].seed_hi << 32) | S[g].seed_lo;
      const U4 r = draw(seed, kStSched, 0, step, k);
      sch[g][k] = uniform_int(r.x, 6) | ((r.y >> 31) << 3);
    }
    if (si > 0) {
      // the ghosts: the neighbours' boundary planes at the end of the last step
      if (tid == 0) {
        const uint64_t want = ring_version(step - 1, kRingHalf - 1);
        for (uint32_t g = 0; g < nsl && !abort_s; ++g)
          if (S[g].ns > 1 && (!wait_version(ver + 2 * S[g].above, want, status) ||
                              !wait_version(ver + 2 * S[g].below + 1, want, status)))
            abort_s = 1;
      }
      __syncthreads();
      const uint32_t prev = (uint32_t)((step - 1) & 1u) * kRingHalf + kRingHalf - 1;
      for (uint32_t g = 0; g < nsl && !abort_s; ++g) {
        const RSlab& s = S[g];
        if (s.ns == 1) continue;
        const uint32_t LL = s.L * s.L;
        const uint32_t* hi = rings + s.ring_hi + prev * LL;
        const uint32_t* lo = rings + s.ring_lo + (2 * kRingHalf + prev) * LL;
        uint32_t* D = lds + s.lds;
        for (uint32_t i = tid; i < 2 * LL; i += nt)
          D[i < LL ? (s.T + 1) * LL + i : i - LL] = ring_load(i < LL ? hi + i : lo + (i - LL));
      }
    }
    __syncthreads();
    // per slab: ring slots of this step's half-sweeps and the items of every
    // phase
    if (tid < nsl) {
      const RSlab& s = S[tid];
      uint32_t n_dn = 0, n_up = 0, n_hi = 0, n_lo = 0;  // publishes so far: mine (bottom, top), above's, below's
      const uint32_t p_bot = parity_of((int)s.z0, s.L), p_top = parity_of((int)(s.z0 + s.T - 1), s.L);
      const uint32_t p_abv = parity_of((int)(s.z0 + s.T), s.L), p_blw = parity_of((int)s.z0 - 1, s.L);
      const uint32_t LL = s.L * s.L, qL = s.L / 4;
      for (uint32_t fam = 0; fam < 2; ++fam) {
        // cube origins: planes [-1, T-1] (ghost below included) of the
        // family's parity; [0, T-1] for a whole-lattice window
        const int zlo = s.ns > 1 ? -1 : 0;
        const int first = zlo + (parity_of((int)s.z0 + zlo, s.L) == fam ? 0 : 1);
        seg[fam][tid] = Seg{first, first <= (int)s.T - 1 ? (uint32_t)((int)s.T - 1 - first) / 2 + 1 : 0u,
                            qL * qL, s.mQQ, 2};
      }
      for (int k = 0; k < 12; ++k) {
        const uint32_t arm = sch[tid][k] & 7u, q = (sch[tid][k] >> 3) ? 0u : 1u;
        uint8_t r = 0, w = 0;
        if (s.ns > 1 && arm == 4) {
          if (q == p_top) {
            r = (uint8_t)n_hi;  // 0: nothing published above in this step yet
            w = (uint8_t)(++n_up);
          }
        } else if (s.ns > 1 && arm == 5) {
          if (q == p_bot) {
            r = (uint8_t)n_lo;
            w = (uint8_t)(++n_dn);
          }
        }
        // the neighbours publish on the same schedule (one state point)
        if (s.ns > 1 && arm == 5 && q == p_abv) ++n_hi;
        if (s.ns > 1 && arm == 4 && q == p_blw) ++n_lo;
        rd[tid][k] = r;
        wr[tid][k] = w;
        if (arm < 4) {
          seg[2 + k][tid] = Seg{0, s.T, LL / 8, s.mP8, 1};
        } else {
          const int first = parity_of((int)s.z0, s.L) == q ? 0 : 1;
          seg[2 + k][tid] = Seg{first, first < (int)s.T ? (s.T - (uint32_t)first + 1) / 2 : 0u, LL / 4, s.mP4, 2};
        }
      }
    }
    __syncthreads();
    if (tid < kPhases) {
      // items are handed out in chunks of 64 (one per wave), each slab's
      // items padded to whole chunks, so that a wave's slab is uniform
      uint32_t a = 0;
      for (uint32_t g = 0; g < nsl; ++g) {
        pre[tid][g] = a;
        a += (seg[tid][g].np * seg[tid][g].Q + 63) / 64;
      }
      pre[tid][nsl] = a;
    }
    __syncthreads();
    const uint32_t lane = tid & 63u, wave = uni(tid >> 6), nwave = nt >> 6;
    // The waves take the phase's 64-item chunks in turn (wave w: chunks w,
    // w + nwave, ...: consecutive chunks on different SIMDs).  A wave keeps
    // the scalar fields of its chunk's slab until a chunk lies past that
    // slab.  body(c, pi, u) handles the lane's item: plane pi of the slab's
    // phase segment, in-plane item u.
    auto walk = [&](uint32_t ph, auto&& body) {
      const uint32_t C = uni(pre[ph][nsl]);
      Cur c;
      uint32_t sbeg = 0, send = 0, cnt = 0;
      for (uint32_t ch = wave; ch < C && !abort_s; ch += nwave) {
        if (ch >= send) {
          c = chunk_slab(pre[ph], seg[ph], S, ch);
          sbeg = uni(pre[ph][c.g]);
          send = uni(pre[ph][c.g + 1]);
          cnt = c.sg.np * c.sg.Q;
        }
        const uint32_t loc = (ch - sbeg) * 64 + lane;
        if (loc < cnt) {
          const uint32_t pi = fdiv(loc, FastDiv{c.sg.Q, c.sg.mQ});
          body(c, pi, loc - pi * c.sg.Q);
        }
      }
    };
    // ---------------------------------------------------------- cube sweep
    for (uint32_t fam = 0; fam < 2; ++fam) {
      walk(fam, [&](const Cur& c, uint32_t pi, uint32_t u) {
        const uint32_t g = c.g;
        const uint32_t jj = fdiv(u, FastDiv{c.qL, c.mQ}), ii = u - jj * c.qL;
        const int zr = c.sg.first + 2 * (int)pi;
        const uint32_t ox = 2 * ii + fam, oy = 2 * jj + fam;
        const int zg = (int)c.z0 + zr;
        const uint32_t zgw = (uint32_t)(zg < 0 ? zg + (int)c.L : zg);
        // cube index x/2 + (L/2)(y/2) + (L^2/4) z (gpu_water3D.cu:3405)
        const uint32_t cube_ref = ii + c.hL * jj + (c.LL / 4) * zgw;
        const U4 r = draw(c.seed, kStEta, cube_ref, step, 0);
        const uint32_t wb = c.ns > 1 ? (uint32_t)(zr + 1) : (uint32_t)zr;  // window plane of the origin
        const uint32_t wt = c.ns > 1 ? wb + 1 : (wb + 1 == c.T ? 0u : wb + 1);
        uint32_t* B = lds + c.lds + wb * c.LL;
        uint32_t* U = lds + c.lds + wt * c.LL;
        eta_cube_planes(B, U, thr_e[g], etab, c.L, ox, oy, r.x);
        eta_cube_planes(B, U, thr_e[g], etab, c.L, ox, oy + c.hL, r.y);
        eta_cube_planes(B, U, thr_e[g], etab, c.L, ox + c.hL, oy, r.z);
        eta_cube_planes(B, U, thr_e[g], etab, c.L, ox + c.hL, oy + c.hL, r.w);
      });
      __syncthreads();
    }
    // ----------------------------------------------------- 12 half-sweeps
    for (uint32_t k = 0; k < 12; ++k) {
      const uint32_t ph = 2 + k;
      bool wrote_ring = false;
      walk(ph, [&](const Cur& c, uint32_t pi, uint32_t u) {
        const uint32_t g = c.g;
        const uint32_t L = c.L, hL = c.hL, qL = c.qL, LL = c.LL;
        const uint32_t sc = uni(sch[g][k]);
        const int arm = (int)(sc & 7u);
        const uint32_t q = (sc >> 3) ? 0u : 1u;
        const int ax = arm >> 1;
        const int zr = c.sg.first + (int)(c.sg.step * pi);
        const uint32_t zg = c.z0 + (uint32_t)zr;
        const uint32_t wz = c.ns > 1 ? (uint32_t)(zr + 1) : (uint32_t)zr;  // window plane
        const uint32_t base = c.lds + wz * LL;
        uint32_t x, y, cell[4], nb[4];
        const uint32_t* RG = nullptr;  // the facing neighbours' ring plane, or none (LDS)
        if (ax < 2) {
          if (ax == 0) {  // u = hx + qL y
            y = fdiv(u, FastDiv{qL, c.mQ});
            x = 2 * (u - y * qL) + q;
          } else {  // u = x + hL hy
            const uint32_t hy = fdiv(u, FastDiv{hL, c.mH});
            x = u - hy * hL;
            y = 2 * hy + q;
          }
          quad_plane_cells(base, L, x, y, arm, ax, cell, nb);
        } else {
          y = fdiv(u, FastDiv{hL, c.mH});
          x = u - y * hL;
          const uint32_t b0 = base + x + L * y, dy = hL * L;
          cell[0] = b0;
          cell[1] = b0 + dy;
          cell[2] = b0 + hL;
          cell[3] = b0 + hL + dy;
          const bool edge = arm == 4 ? zr == (int)c.T - 1 : zr == 0;
          const uint32_t r = edge ? uni(rd[g][k]) : 0u;
          if (r) {
            // the neighbour's plane as published after its r-th z half-sweep
            // of this step, read from its ring
            const RSlab& s = S[g];
            const uint32_t nslab = arm == 4 ? (uint32_t)s.above : (uint32_t)s.below;
            if (!wait_version(ver + 2 * nslab + (arm == 4 ? 0 : 1), ring_version(step, r - 1), status)) {
              abort_s = 1;
              return;
            }
            RG = arm == 4 ? rings + s.ring_hi + (half + r - 1) * LL
                          : rings + s.ring_lo + (2 * kRingHalf + half + r - 1) * LL;
            const uint32_t o = x + L * y;
            nb[0] = o;
            nb[1] = o + dy;
            nb[2] = o + hL;
            nb[3] = o + hL + dy;
          } else {
            // neighbour plane in the window (a whole-lattice window wraps)
            const uint32_t nz = arm == 4 ? (c.ns == 1 && wz + 1 == c.T ? 0u : wz + 1)
                                         : (c.ns == 1 && wz == 0 ? c.T - 1 : wz - 1);
            const uint32_t nbase = c.lds + nz * LL;
#pragma unroll
            for (int m = 0; m < 4; ++m) nb[m] = nbase + (cell[m] - base);
          }
        }
        uint32_t w[4], wn[4];
        if (RG) {
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            w[m] = lds[cell[m]];
            wn[m] = ring_load(RG + nb[m]);
          }
        } else {
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            w[m] = lds[cell[m]];
            wn[m] = lds[nb[m]];
          }
        }
        const U4 rq = draw(c.seed, kStSigma, x + L * y + LL * zg, step, k);
        sigma_quad_regs(w, wn, thr_s[g], arm, rq);
#pragma unroll
        for (int m = 0; m < 4; ++m) lds[cell[m]] = w[m];
        const uint32_t wslot = uni(wr[g][k]);
        if (ax == 2 && (arm == 4 ? zr == (int)c.T - 1 : zr == 0) && wslot) {
          // a boundary plane all of whose cells move in this half-sweep: the
          // new words go to this slab's ring as well
          uint32_t* R = rings + uni(S[g].ring) + ((arm == 4 ? 2 * kRingHalf : 0u) + half + wslot - 1) * LL;
#pragma unroll
          for (int m = 0; m < 4; ++m) ring_store(R + (cell[m] - base), w[m]);
          wrote_ring = true;
        }
      });
      if (wrote_ring) ring_stores_done();
      __syncthreads();
      if (tid < nsl && wr[tid][k] && !abort_s) {
        const uint32_t arm = sch[tid][k] & 7u;
        publish(ver + 2 * (G.first + tid) + (arm == 4 ? 1 : 0), ring_version(step, wr[tid][k] - 1));
      }
    }
    // ------------------------------------- boundary planes at step end
    bool copied = false;
    for (uint32_t g = 0; g < nsl && !abort_s; ++g) {
      const RSlab& s = S[g];
      if (s.ns == 1) continue;
      const uint32_t LL = s.L * s.L;
      uint32_t* dn = rings + s.ring + (half + kRingHalf - 1) * LL;
      uint32_t* up = rings + s.ring + (2 * kRingHalf + half + kRingHalf - 1) * LL;
      const uint32_t* bot = lds + s.lds + LL;
      const uint32_t* top = lds + s.lds + s.T * LL;
      for (uint32_t i = tid; i < 2 * LL; i += nt) {
        if (i < LL)
          ring_store(dn + i, bot[i]);
        else
          ring_store(up + (i - LL), top[i - LL]);
        copied = true;
      }
    }
    if (copied) ring_stores_done();
    __syncthreads();
    if (tid < nsl && S[tid].ns > 1 && !abort_s) {
      publish(ver + 2 * (G.first + tid), ring_version(step, kRingHalf - 1));
      publish(ver + 2 * (G.first + tid) + 1, ring_version(step, kRingHalf - 1));
    }
  }
  // the neighbours have read their windows (they published after staging)
  // before this workgroup overwrites its planes in `words`
  if (tid == 0 && nsteps > 0) {
    const uint64_t want = ring_version(step0 + nsteps - 1, kRingHalf - 1);
    for (uint32_t g = 0; g < nsl && !abort_s; ++g)
      if (S[g].ns > 1 && (!wait_version(ver + 2 * S[g].above, want, status) ||
                          !wait_version(ver + 2 * S[g].below + 1, want, status)))
        abort_s = 1;
  }
  __syncthreads();
  if (abort_s) return;  // after the last barrier
  for (uint32_t g = 0; g < nsl; ++g) {
    const RSlab& s = S[g];
    const uint32_t LL = s.L * s.L, Q4 = LL / 4;
    uint4* W4 = reinterpret_cast<uint4*>(words + s.woff + s.z0 * LL);
    const uint4* Wl = reinterpret_cast<const uint4*>(lds + s.lds + (s.ns > 1 ? LL : 0u));
    for (uint32_t i = tid; i < s.T * Q4; i += nt) W4[i] = Wl[i];
  }
}

uint32_t resident_lds_limit() { return 160 * 1024 - kResidentStatic; }

static void set_resident_attribute() {
  // the dynamic-LDS limit is a per-device function attribute
  static std::atomic<uint64_t> attr_set{0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t bit = 1ull << (dev & 63);
  if (!(attr_set.load(std::memory_order_acquire) & bit)) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_resident), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)resident_lds_limit());
    attr_set.fetch_or(bit, std::memory_order_acq_rel);
  }
}

hipError_t resident_occupancy(int* per_cu, uint32_t block, uint32_t lds_bytes) {
  set_resident_attribute();
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_resident, (int)block, (size_t)lds_bytes);
}

void launch_resident(hipStream_t st, const RGroupEnt* grp, uint32_t ngroups, uint32_t block, uint32_t lds_bytes,
                     const RSlab* slabs, const PointDev* pts, uint32_t* words, uint32_t* rings, uint64_t* ver,
                     const uint64_t* stepp, uint32_t step_off, uint32_t nsteps, uint32_t* status) {
  set_resident_attribute();
  hipEvent_t t0, t1;
  get_launch_timing(&t0, &t1);
  if (t0)
    hipExtLaunchKernelGGL(k_resident, dim3(ngroups), dim3(block), lds_bytes, st, t0, t1, 0, grp, slabs, pts, words,
                          rings, ver, stepp, step_off, nsteps, status);
  else
    hipLaunchKernelGGL(k_resident, dim3(ngroups), dim3(block), lds_bytes, st, grp, slabs, pts, words, rings, ver,
                       stepp, step_off, nsteps, status);
}

}  // namespace cvf
