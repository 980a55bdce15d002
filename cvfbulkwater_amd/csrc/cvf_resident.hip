// Resident Metropolis kernel: the lattice stays in LDS for a whole sequence
// of MC steps (the reference schedule: eta cube sweep + 12 sigma
// half-sweeps per step, gpu_water3D.cu:1346-1417, 1890-2008, 3374-3466).
//
// Each state point is cut into ns z-slabs (ns = 1: the whole lattice in one
// window, periodic in LDS); slabs of one point may differ in thickness.
// The host fills the workgroups with slabs, about the same number of cells
// each, at most one or two workgroups per CU so that every workgroup of the
// launch is resident at once, and gives every slab of a workgroup its own
// range of waves, in proportion to its cells.  A wave works on its slab
// only: the slab's fields sit in scalar registers for the whole launch, and
// the items of every phase are walked without a division.  A workgroup
// stages its slabs once, runs `nsteps` MC steps on them in LDS and writes
// them back once.  A split slab's window holds its T planes plus a ghost
// copy of the plane below (window plane 0) and of the plane above (window
// plane T + 1), owned by the neighbouring slabs.
//
// Window planes are stored by quads: quad (x, y), x, y < L/2, is the four
// cells (x, y), (x, y + L/2), (x + L/2, y), (x + L/2, y + L/2), which share
// one Philox block in every sigma half-sweep and move together; they are
// words 4 (y L/2 + x) .. + 3 of the plane, in that order (half h = 2 [x
// half] + [y half]).  A quad is one 16-byte LDS load and store, and a cube
// item (four cubes L/2 apart in x and y) reads and writes eight whole
// quads.  `words` keeps the reference's x + L y + L^2 z order; staging and
// write-back convert, and the ring planes carry the window layout.
//
// Only z half-sweeps and the cube sweep read across a slab boundary:
//   * a +z half-sweep (arm 4) of the top plane reads sigma_5 of the plane
//     above, which changes only in -z half-sweeps of that plane, and a -z
//     half-sweep of the bottom plane reads sigma_4 of the plane below;
//   * a cube whose origin plane is the last plane of a slab (or the ghost
//     below) spans the boundary; both slabs compute it, from the same words
//     and the same cube-keyed random numbers, so both copies agree.
// A slab therefore publishes its bottom (top) plane after every -z (+z)
// half-sweep in which that plane is active and a neighbour reads it before
// it moves again, and both boundary planes at the end of every step (a
// one-plane slab its plane once, in the lower half), into rings in global
// memory: slots (step & 1) * 32 + 2k, + 1 after half-sweep
// k, slots (step & 1) * 32 + 30, 31 at the step end.  Every ring word is 64
// bits, a cell word and a tag of the step; a reader loads the words it
// needs and loads a word again while its tag is older, so neither a
// writer's store completion nor a version round trip lies on the path from
// one slab to the next.  A slot's previous contents (two steps back, or
// from before the rings were cleared) never carry the awaited tag.  Only
// the launch's last step also publishes a version, which the neighbours
// await before they write their planes back to `words`.
//
// Coherence across the XCDs' L2 caches: every ring word and every version
// is accessed with agent-scope relaxed atomics (single-location coherent
// loads and stores that bypass the XCDs' non-coherent L2 caching), so no
// L2 writeback or invalidation is needed.  A tagged word needs no order:
// its tag and cell word are one 64-bit store.  The last step's version is
// stored after the workgroup barrier that follows the completion of the
// slab's ring stores (s_waitcnt).
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

// The version a slab publishes after the last step of a launch.
__device__ __forceinline__ uint64_t end_version(uint64_t step) { return step + 1; }

// Step-end planes: 64-bit ring words, the cell word in the low half and
// end_tag(step) in the high half, in ring slots kEndSlot, kEndSlot + 1 of
// the step's half.  The tag of a slot's previous contents (two steps back,
// or a stale value from before the rings were cleared) is never the tag a
// reader waits for.
constexpr uint32_t kEndSlot = kRingHalf - 2;
// first ring slot (of two) of the publish after half-sweep k
__device__ __forceinline__ uint32_t ring_slot(uint32_t k) { return 2 * k; }
__device__ __forceinline__ uint32_t end_tag(uint64_t step) { return (uint32_t)(step + 1) | 0x80000000u; }
__device__ __forceinline__ uint64_t ring_load64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ring_store64(uint64_t* p, uint64_t v) {
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

// Load each of a lane's N words p[m] again until its tag (high half) is
// `tag`, starting from the values x[m] already loaded, in one loop of the
// whole wave: every pass checks all the words of every active lane, leaves
// when none is stale and reloads only the stale ones.  So every load is
// consumed on every path, and the compiler needs no wait for a load that
// might still be in flight later (on this GPU loads and stores share one
// counter, and such a wait would also wait out the ring stores).  False
// on time-out or abort (then some x[m] are stale).
template <int N>
__device__ __forceinline__ bool wait_tags(const uint64_t* const (&p)[N], uint32_t tag, uint64_t (&x)[N],
                                          uint32_t* status) {
  for (uint32_t it = 0;; ++it) {
    bool stale[N], any = false;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      stale[m] = (uint32_t)(x[m] >> 32) != tag;
      any = any || stale[m];
    }
    if (!__ballot(any)) return true;
    if (((it & 63u) == 63u && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ||
        it >= kSpinLimit) {
      __hip_atomic_fetch_or(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int m = 0; m < N; ++m)
      if (stale[m]) x[m] = ring_load64(p[m]);
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
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Chessboard sigma rule on one word w (facing neighbour word wn, random
// word rw): the cell's new word, as sigma_quad_regs computes each of four.
__device__ __forceinline__ uint32_t sigma_cell(uint32_t w, uint32_t wn, const uint32_t* thr, int arm, uint32_t rw) {
  uint32_t sp, ra;
  split_word(rw, 6, sp, ra);
  const uint32_t sa = sigma_of(w, arm);
  const uint32_t t = thr[sigma_thr_index(w, wn, arm, sa, sp)];
  const uint32_t nw = (w & ~(7u << (3 * arm))) | (sp << (3 * arm));
  return sp != sa && ra < t ? nw : w;
}

// A phase runs fine (one cell or cube per lane) when its cells fit the
// slab's lanes and take at most 512 of them, i.e. at most two waves per
// SIMD: beyond that the longer per-quad chain on fewer waves is cheaper.
__device__ __forceinline__ bool fine_fits(uint32_t cells, uint32_t lanes) { return cells <= lanes && cells <= 512; }

// Word h of a Philox block (h lane-dependent).
__device__ __forceinline__ uint32_t pick4(const U4& r, uint32_t h) {
  return h == 0 ? r.x : h == 1 ? r.y : h == 2 ? r.z : r.w;
}

// Chessboard sigma rule on four words held in registers (facing neighbour
// words wn, one Philox block): the cells' new words.  The four threshold
// loads are issued together (the index is in range also when the proposal
// equals the current value), then the four decisions are made.
__device__ __forceinline__ void sigma_quad_regs(uint32_t (&w)[4], const uint32_t (&wn)[4], const uint32_t* thr,
                                                int arm, const U4& q) {
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
  for (int m = 0; m < 4; ++m) {
    const uint32_t nw = (w[m] & ~(7u << (3 * arm))) | (sp[m] << (3 * arm));
    w[m] = act[m] && ra[m] < t[m] ? nw : w[m];
  }
}

// The draws of one phase, draw(seed, stream, site, step, sub) for every
// item's site: the part that does not depend on the site, once per phase
// (in scalar registers) ...
__device__ __forceinline__ PhiloxFixed phase_draws(uint64_t seed, uint32_t stream, uint64_t step, uint32_t sub) {
  return philox_fixed((uint32_t)step, (stream << 24) | sub, (uint32_t)(step >> 32), (uint32_t)seed,
                      (uint32_t)(seed >> 32));
}
// ... and one item's draw.  The round-2 key goes through an empty asm
// statement, so that the compiler rebuilds the later round keys with
// scalar adds in each iteration instead of holding all of them in scalar
// registers across the loop (which spills the phase's other scalars).
__device__ __forceinline__ U4 draw_site(uint32_t site, const PhiloxFixed& f) {
  uint32_t k0 = f.k0, k1 = f.k1;
  asm volatile("" : "+s"(k0), "+s"(k1));
  return philox_site(site, f, k0, k1);
}

// A lane's walk over the items of one phase of its slab: item it = pi * Q +
// u, starting at (wave - w0) * 64 + lane and advancing by the slab's
// 64 nw lanes (dp planes and du items), without a division.
struct Walk {
  uint32_t pi, u;
};
__device__ __forceinline__ void advance(Walk& w, uint32_t Q, uint32_t dp, uint32_t du) {
  w.u += du;
  w.pi += dp;
  if (w.u >= Q) {
    w.u -= Q;
    ++w.pi;
  }
}
__device__ __forceinline__ Walk walk_start(uint32_t it, uint32_t Q, uint32_t mQ) {
  const uint32_t pi = fdiv(it, FastDiv{Q, mQ});
  return Walk{pi, it - pi * Q};
}

// (a & m) | (b & ~m) in one instruction
__device__ __forceinline__ uint32_t bitsel(uint32_t m, uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
  return r;
}

// The four words of a quad (16-byte aligned window index w).
__device__ __forceinline__ void quad_load(const uint32_t* lds, uint32_t w, uint32_t (&v)[4]) {
  const uint4 t = *reinterpret_cast<const uint4*>(lds + w);
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ __forceinline__ void quad_store(uint32_t* lds, uint32_t w, const uint32_t (&v)[4]) {
  *reinterpret_cast<uint4*>(lds + w) = make_uint4(v[0], v[1], v[2], v[3]);
}

// One in-plane half-sweep (arm along axis AX: 0 x, 1 y) of a slab's T
// planes: the lane's quads from item li on, 64 nw items apart.  Item u of
// a plane is the quad with origin (2 (u mod L/4) + q, u div L/4) (x arm)
// or (u mod L/2, 2 (u div L/2) + q) (y arm); its index in the plane is
// 2u + q, resp. u + (L/2)(u div L/2 + q).  The facing quad is the next
// one along the axis (the previous one for a - arm); across the periodic
// edge it is the quad at the other end, with the halves along the axis
// swapped.  WRAP: the phase's parity holds the origins next to the edge
// (without it no item wraps, and the test and the selects are left out).
template <int AX, bool WRAP>
__device__ __forceinline__ void inplane_items(uint32_t* lds, uint32_t li, uint32_t L, uint32_t T, uint32_t z0,
                                              uint32_t pbase, int arm, uint32_t q, uint32_t edge, int dn, int dw,
                                              const PhiloxFixed& pf, FastDiv dH, FastDiv dQ, uint32_t mP8,
                                              uint32_t dp8, uint32_t du8, const uint32_t* thr) {
  const uint32_t hL = L / 2, qL = L / 4, LL = L * L, Q8 = LL / 8;
  for (Walk it = walk_start(li, Q8, mP8); it.pi < T; advance(it, Q8, dp8, du8)) {
    uint32_t x, y, quad;
    if (AX == 0) {
      y = fdiv(it.u, dQ);
      x = 2 * (it.u - y * qL) + q;
      quad = 2 * it.u + q;
    } else {
      const uint32_t hy = fdiv(it.u, dH);
      x = it.u - hy * hL;
      y = 2 * hy + q;
      quad = it.u + hL * (hy + q);
    }
    const bool wrap = WRAP && (AX == 0 ? x : y) == edge;
    const uint32_t cw = pbase + it.pi * LL + 4 * quad;
    const uint32_t nw = cw + 4u * (uint32_t)(wrap ? dw : dn);
    uint32_t w[4], wn[4];
    quad_load(lds, cw, w);
    quad_load(lds, nw, wn);
    if (WRAP) {
      // swap the halves along the axis where the facing quad wraps (a bit
      // select, so that the four selects stay four instructions)
      const uint32_t sel = wrap ? ~0u : 0u;
      uint32_t nv[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) nv[m] = wn[m];
#pragma unroll
      for (int m = 0; m < 4; ++m) wn[m] = bitsel(sel, nv[m ^ (AX == 0 ? 2 : 1)], nv[m]);
    }
    const U4 rq = draw_site(x + L * y + LL * (z0 + it.pi), pf);
    sigma_quad_regs(w, wn, thr, arm, rq);
    quad_store(lds, cw, w);
  }
}

template <int AX, bool FINE>
__device__ __forceinline__ void sweep_inplane(uint32_t* lds, uint32_t li, uint32_t L, uint32_t T, uint32_t z0,
                                              uint32_t pbase, uint64_t seed, uint64_t step, uint32_t k, int arm,
                                              uint32_t q, FastDiv dH, FastDiv dQ, uint32_t mP8, uint32_t dp8,
                                              uint32_t du8, const uint32_t* thr) {
  const uint32_t hL = L / 2, qL = L / 4, LL = L * L, Q8 = LL / 8;
  const bool up = (arm & 1) == 0;
  const uint32_t st = AX == 0 ? 1u : hL;                    // quads per step along the axis
  const uint32_t edge = up ? hL - 1 : 0u;                   // origin coordinate whose neighbour wraps
  const int dn = up ? (int)st : -(int)st;                   // facing quad, no wrap
  const int dw = up ? -(int)((hL - 1) * st) : (int)((hL - 1) * st);  // facing quad across the edge
  const PhiloxFixed pf = phase_draws(seed, kStSigma, step, k);
  const uint32_t lanes = dp8 * Q8 + du8;
  if (FINE && fine_fits(4 * T * Q8, lanes) && 2 * T * Q8 < 128) {
    // a phase of fewer than 64 quads: one cell per lane (cell li mod 4 of
    // quad li div 4), so that it still spreads over two waves
    const uint32_t h = li & 3u;
    const Walk it = walk_start(li >> 2, Q8, mP8);
    if (it.pi < T) {
      uint32_t x, y, quad;
      if (AX == 0) {
        y = fdiv(it.u, dQ);
        x = 2 * (it.u - y * qL) + q;
        quad = 2 * it.u + q;
      } else {
        const uint32_t hy = fdiv(it.u, dH);
        x = it.u - hy * hL;
        y = 2 * hy + q;
        quad = it.u + hL * (hy + q);
      }
      const bool wrap = (AX == 0 ? x : y) == edge;
      const uint32_t cw = pbase + it.pi * LL + 4 * quad;
      const uint32_t nw = cw + 4u * (uint32_t)(wrap ? dw : dn) + (wrap ? h ^ (AX == 0 ? 2u : 1u) : h);
      const uint32_t w = lds[cw + h], wn = lds[nw];
      const U4 rq = draw_site(x + L * y + LL * (z0 + it.pi), pf);
      lds[cw + h] = sigma_cell(w, wn, thr, arm, pick4(rq, h));
    }
    return;
  }
  if (FINE && fine_fits(2 * T * Q8, lanes)) {
    // pairs: lane li takes cells h0 = 2 (li mod 2) and h0 + 1 of quad
    // li div 2, so that a thin slab's phase is a short dependent chain per
    // lane on half the waves one cell per lane would take (four lanes
    // recomputing a quad's Philox block cost more issue than a second cell's
    // chain).  The facing cells of the pair are
    // the pair at h0 of the facing quad, across the edge the pair at h0 ^ 2
    // (x arm) or the same pair in swapped order (y arm).
    const uint32_t h0 = 2 * (li & 1u);
    const Walk it = walk_start(li >> 1, Q8, mP8);
    if (it.pi < T) {
      uint32_t x, y, quad;
      if (AX == 0) {
        y = fdiv(it.u, dQ);
        x = 2 * (it.u - y * qL) + q;
        quad = 2 * it.u + q;
      } else {
        const uint32_t hy = fdiv(it.u, dH);
        x = it.u - hy * hL;
        y = 2 * hy + q;
        quad = it.u + hL * (hy + q);
      }
      const bool wrap = (AX == 0 ? x : y) == edge;
      const uint32_t cw = pbase + it.pi * LL + 4 * quad + h0;
      const uint32_t nw = cw - h0 + 4u * (uint32_t)(wrap ? dw : dn) + (AX == 0 && wrap ? h0 ^ 2u : h0);
      const uint2 w = *reinterpret_cast<const uint2*>(lds + cw);
      const uint2 n = *reinterpret_cast<const uint2*>(lds + nw);
      const bool sw = AX == 1 && wrap;
      const U4 rq = draw_site(x + L * y + LL * (z0 + it.pi), pf);
      const uint32_t r0 = h0 ? rq.z : rq.x, r1 = h0 ? rq.w : rq.y;
      const uint32_t a = sigma_cell(w.x, sw ? n.y : n.x, thr, arm, r0);
      const uint32_t b = sigma_cell(w.y, sw ? n.x : n.y, thr, arm, r1);
      *reinterpret_cast<uint2*>(lds + cw) = make_uint2(a, b);
    }
    return;
  }
  // L/2 is even, so the origins along the axis have parity q and the wrapping
  // origin has parity 1 (+ arm, L/2 - 1) or 0 (- arm, 0): a phase of the
  // other parity never reaches across the edge (a wave-uniform choice)
  if (q == (up ? 1u : 0u))
    inplane_items<AX, true>(lds, li, L, T, z0, pbase, arm, q, edge, dn, dw, pf, dH, dQ, mP8, dp8, du8, thr);
  else
    inplane_items<AX, false>(lds, li, L, T, z0, pbase, arm, q, edge, dn, dw, pf, dH, dQ, mP8, dp8, du8, thr);
}

// Cube moves on the four cubes of a cube item in the quad layout: origins
// (ox + a L/2, oy + b L/2), ox, oy < L/2, on bottom plane B and top plane U
// (window indices), random words r.x .. r.w in the order (a, b) = (0,0),
// (0,1), (1,0), (1,1), i.e. cube c = 2a + b takes half c of each corner
// quad; a corner across the periodic edge (ox or oy = L/2 - 1) lies in
// quad 0 along that axis with the halves swapped.  The four cubes are
// disjoint and own all four halves of each of their eight corner quads, so
// the quads are loaded before the first move and stored back whole after
// the last.
__device__ __forceinline__ void swap_halves(uint32_t (&v)[4], uint32_t sel, int bit) {
  uint32_t t[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) t[h] = bitsel(sel, v[h ^ bit], v[h]);
#pragma unroll
  for (int h = 0; h < 4; ++h) v[h] = t[h];
}
// Cube c alone of the cube item eta_cube_item_q describes (random word
// rw = word c of the item's block): its eight corner words are half
// c ^ (the corner's swap) of the eight corner quads.
__device__ __forceinline__ void eta_cube_one_q(uint32_t* lds, uint32_t B, uint32_t U, const uint64_t* thr,
                                               const uint64_t* etab, uint32_t hL, uint32_t ox, uint32_t oy, uint32_t c,
                                               uint32_t rw) {
  const bool ex = ox + 1 == hL, ey = oy + 1 == hL;
  const uint32_t x1 = ex ? 0u : ox + 1, y1 = ey ? 0u : oy + 1;
  uint32_t idx[8], w[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint32_t f = ((b & 1) && ex ? 2u : 0u) | ((b & 2) && ey ? 1u : 0u);
    idx[b] = ((b & 4) ? U : B) + 4 * (((b & 2) ? y1 : oy) * hL + ((b & 1) ? x1 : ox)) + (c ^ f);
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) w[b] = lds[idx[b]];
  eta_cube_move(w, thr, etab, rw, [&](int b, uint32_t v) { lds[idx[b]] = v; });
}

// EDGE: the item may reach across the periodic edge (family 1; a
// family-0 item's origins are even, below L/2 - 1, so it never does).
template <bool EDGE>
__device__ __forceinline__ void eta_cube_item_q(uint32_t* lds, uint32_t B, uint32_t U, const uint64_t* thr,
                                                const uint64_t* etab, uint32_t hL, uint32_t ox, uint32_t oy,
                                                const U4& r) {
  const bool ex = ox + 1 == hL, ey = oy + 1 == hL;
  const uint32_t x1 = ex ? 0u : ox + 1, y1 = ey ? 0u : oy + 1;
  const uint32_t sx = ex ? ~0u : 0u, sy = ey ? ~0u : 0u;
  // corner b = bx + 2 by + 4 bz: quad (bx ? x1 : ox, by ? y1 : oy) of plane bz ? U : B
  uint32_t qw[8], v[8][4];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    qw[b] = ((b & 4) ? U : B) + 4 * (((b & 2) ? y1 : oy) * hL + ((b & 1) ? x1 : ox));
    quad_load(lds, qw[b], v[b]);
  }
  // half c of every corner quad for cube c
  if (EDGE) {
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      if (b & 1) swap_halves(v[b], sx, 2);
      if (b & 2) swap_halves(v[b], sy, 1);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t rw = c == 0 ? r.x : c == 1 ? r.y : c == 2 ? r.z : r.w;
    uint32_t w[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) w[b] = v[b][c];
    eta_cube_move(w, thr, etab, rw, [&](int b, uint32_t nv) { v[b][c] = nv; });
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    if (EDGE && (b & 2)) swap_halves(v[b], sy, 1);
    if (EDGE && (b & 1)) swap_halves(v[b], sx, 2);
    quad_store(lds, qw[b], v[b]);
  }
}

}  // namespace

// grid = one workgroup per RGroupEnt; dynamic LDS = the largest sum of
// window sizes of one workgroup.  The slab records sit in LDS and each
// phase reads the fields it uses (wave-uniform), so that they are not all
// held in scalar registers through the step loop.
template <bool FINE>
__global__ __launch_bounds__(1024) void k_resident(const RGroupEnt* __restrict__ grp, const RSlab* __restrict__ slabs,
                                                    const PointDev* __restrict__ pts, uint32_t* __restrict__ words,
                                                    uint32_t* __restrict__ rings, uint64_t* __restrict__ ver,
                                                    const uint64_t* __restrict__ stepp, uint32_t step_off,
                                                    uint32_t nsteps, uint32_t* __restrict__ status) {
  extern __shared__ __align__(16) uint32_t lds[];
  __shared__ RSlab Ssh[kMaxRSlabs];
  __shared__ uint32_t thr_s[kMaxRSlabs][33];
  __shared__ uint64_t thr_e[kMaxRSlabs][9];
  __shared__ uint64_t etab[9];
  // the step's half-sweeps of each slab: arm | parity << 3 | rd << 4 |
  // wr << 8 (rd, wr: ring slots, see below)
  __shared__ uint32_t sch_s[kMaxRSlabs][12];
  __shared__ uint32_t abort_s;
  // per slab: the slab below is one plane thick (its step-end plane then
  // sits in the lower half of its ring only)
  __shared__ uint32_t below1_s[kMaxRSlabs];
  const RGroupEnt G = grp[blockIdx.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = uni(tid >> 6);
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(slabs + G.first);
    uint32_t* dst = reinterpret_cast<uint32_t*>(Ssh);
    for (uint32_t i = tid; i < G.count * (uint32_t)(sizeof(RSlab) / 4); i += blockDim.x) dst[i] = src[i];
  }
  if (tid == 0) abort_s = 0;
  if (tid < 9) etab[tid] = eta_table_entry(tid);
  __syncthreads();
  // this wave's slab: the one whose wave range holds it (none: the wave
  // only keeps the workgroup's barriers)
  uint32_t gs = kMaxRSlabs;
  for (uint32_t g = 0; g < G.count; ++g)
    if (wave >= Ssh[g].w0 && wave < Ssh[g].w0 + Ssh[g].nw) gs = g;
  gs = uni(gs);
  const bool has = gs < kMaxRSlabs;
  const RSlab& S = Ssh[has ? gs : 0];
#define F(f) uni((uint32_t)S.f)
  const uint32_t li = has ? (wave - F(w0)) * 64u + lane : 0u;  // the lane's first item
  if (has && wave == F(w0)) {
    const PointDev& P = pts[F(point)];
    if (lane < 33) thr_s[gs][lane] = sigma_thr32(P.thr_sigma[lane]);
    if (lane < 9) thr_e[gs][lane] = P.thr_eta[lane];
    if (lane == 0) below1_s[gs] = F(ns) > 1 && slabs[S.below].T == 1 ? 1u : 0u;
  }
  if (has) {
    // the slab's window, staged by its waves one quad (16 bytes of LDS,
    // four words of `words`) per lane and item
    const uint32_t L = F(L), hL = L / 2, LL = L * L, Q4 = LL / 4, gofs = F(ns) > 1 ? 1u : 0u,
                   np = F(T) + 2 * gofs;
    const uint32_t* Wp = words + F(woff);
    uint4* D4 = reinterpret_cast<uint4*>(lds + F(lds));
    const FastDiv dq{Q4, F(mP4)}, dH{hL, F(mH)};
    const int zs = (int)F(z0) - (int)gofs;
    if ((hL & 3u) == 0) {
      // four quads along x per lane and item: 16-byte rows of `words`
      for (uint32_t i = 4 * li; i < np * Q4; i += 256 * F(nw)) {
        const uint32_t p = fdiv(i, dq), r = i - p * Q4, y = fdiv(r, dH), x = r - y * hL;
        const int zg = zs + (int)p;
        const uint32_t z = (uint32_t)(zg < 0 ? zg + (int)L : zg >= (int)L ? zg - (int)L : zg);
        const uint4* s = reinterpret_cast<const uint4*>(Wp + z * LL + x + L * y);
        const uint4 a = s[0], c = s[hL / 4], b = s[hL * L / 4], d = s[(hL * L + hL) / 4];
        D4[i] = make_uint4(a.x, b.x, c.x, d.x);
        D4[i + 1] = make_uint4(a.y, b.y, c.y, d.y);
        D4[i + 2] = make_uint4(a.z, b.z, c.z, d.z);
        D4[i + 3] = make_uint4(a.w, b.w, c.w, d.w);
      }
    } else {
      for (uint32_t i = li; i < np * Q4; i += 64 * F(nw)) {
        const uint32_t p = fdiv(i, dq), r = i - p * Q4, y = fdiv(r, dH), x = r - y * hL;
        const int zg = zs + (int)p;
        const uint32_t z = (uint32_t)(zg < 0 ? zg + (int)L : zg >= (int)L ? zg - (int)L : zg);
        const uint32_t* s = Wp + z * LL + x + L * y;
        D4[i] = make_uint4(s[0], s[hL * L], s[hL], s[hL * L + hL]);
      }
    }
  }
  __syncthreads();
  const uint64_t step0 = *stepp + step_off;
  for (uint32_t si = 0; si < nsteps; ++si) {
    const uint64_t step = step0 + si;
    const uint32_t half = (uint32_t)(step & 1u) * kRingHalf;
    // the step's 12 half-sweeps (lane k: half-sweep k) and their ring slots:
    // rd = 1 + the neighbour's half-sweep whose plane to read (0: the ghost
    // is current), wr = 1 when this slab's plane goes to ring slot k
    if (has) {
      const uint32_t L = F(L), z0 = F(z0), T = F(T);
      const uint64_t seed = ((uint64_t)F(seed_hi) << 32) | F(seed_lo);
      // the slab's first wave draws them and keeps them for every wave of
      // the slab
      if (wave == F(w0)) {
        uint32_t schv = 0;
        if (lane < 12) {
          const U4 r = draw(seed, kStSched, 0, step, lane);
          schv = uniform_int(r.x, 6) | ((r.y >> 31) << 3);
        }
        const uint32_t arm = schv & 7u, q = (schv >> 3) ? 0u : 1u;
        const bool in = lane < 12 && F(ns) > 1;
        const uint32_t p_bot = parity_of((int)z0, L), p_top = parity_of((int)(z0 + T - 1), L);
        const uint32_t p_abv = parity_of((int)(z0 + T), L), p_blw = parity_of((int)z0 - 1, L);
        const uint64_t before = (1ull << lane) - 1;                    // lanes before this one
        const uint64_t after = lane < 63 ? ~((2ull << lane) - 1) : 0ull;  // lanes after it
        // the half-sweeps that move this slab's top (up) and bottom (dn)
        // plane, and those that move the neighbours' facing planes: the
        // slab above's bottom (ab), the slab below's top (bl)
        const uint64_t up = __ballot(in && arm == 4 && q == p_top), dn = __ballot(in && arm == 5 && q == p_bot);
        const uint64_t ab = __ballot(in && arm == 5 && q == p_abv), bl = __ballot(in && arm == 4 && q == p_blw);
        // A half-sweep that moves a boundary plane reads the neighbour's
        // plane as of the neighbour's latest such half-sweep (rd = its index
        // + 1; 0: none yet, the ghost holds it), and stores its own moved
        // plane into ring slot k, as tagged words (wr), only if the neighbour
        // reads it: its next reading half-sweep comes before this plane
        // moves again.
        uint32_t rd = 0, wr = 0;
        const bool mine_up = (up >> lane) & 1, mine_dn = (dn >> lane) & 1;
        if (mine_up || mine_dn) {
          const uint64_t other = mine_up ? ab : bl, self = mine_up ? up : dn;
          if (other & before) rd = 64 - (uint32_t)__builtin_clzll(other & before);
          const uint64_t ra = other & after, ua = self & after;
          if (ra && (!ua || __builtin_ctzll(ra) < __builtin_ctzll(ua))) wr = 1;
        }
        if (lane < 12) sch_s[gs][lane] = schv | (rd << 4) | (wr << 8);
      }
      if (si > 0 && F(ns) > 1) {
        // the ghosts: the neighbours' boundary planes at the end of the last
        // step, tagged words (a word whose tag is not yet the last step's is
        // loaded again until it is; bounded)
        const uint32_t LL = L * L, tag = end_tag(step - 1);
        const uint32_t prev = (uint32_t)((step - 1) & 1u) * kRingHalf + kEndSlot;
        const uint64_t* hi = reinterpret_cast<const uint64_t*>(rings + F(ring_hi) + prev * LL);
        // the slab below's top plane; a one-plane slab stores its plane once,
        // in the lower half
        const uint32_t lo_half = uni(below1_s[gs]) ? 0u : 2 * kRingHalf;
        const uint64_t* lo = reinterpret_cast<const uint64_t*>(rings + F(ring_lo) + (lo_half + prev) * LL);
        uint32_t* D = lds + F(lds);
        // eight loads in flight per lane (an index past the end re-reads
        // the last word and is not stored)
        const uint32_t n2 = 2 * LL, stride = 64 * F(nw);
        bool ok = true;
        for (uint32_t i0 = li; i0 < n2; i0 += 8 * stride) {
          uint64_t v[8];
          const uint64_t* src[8];
#pragma unroll
          for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t i = min(i0 + u * stride, n2 - 1);
            src[u] = i < LL ? hi + i : lo + (i - LL);
            v[u] = ring_load64(src[u]);
          }
          ok = wait_tags(src, tag, v, status) && ok;
#pragma unroll
          for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t i = i0 + u * stride;
            if (i < n2) D[i < LL ? (T + 1) * LL + i : i - LL] = (uint32_t)v[u];
          }
        }
        if (__ballot(!ok)) abort_s = 1;
      }
    }
    __syncthreads();
    const bool run = has && !abort_s;
    // ---------------------------------------------------------- cube sweep
    for (uint32_t fam = 0; fam < 2; ++fam) {
      if (run) {
        const uint32_t L = F(L), z0 = F(z0), T = F(T), ns = F(ns), hL = L / 2, qL = L / 4, LL = L * L, Qc = qL * qL;
        const uint64_t seed = ((uint64_t)F(seed_hi) << 32) | F(seed_lo);
        const uint32_t lbase = F(lds), gofs = ns > 1 ? 1u : 0u, dpc = F(dpc), duc = F(duc);
        const FastDiv dQ{qL, F(mQ)};
        const uint64_t* thre = thr_e[gs];
        // origin planes of the family's parity among [-1, T-1] (the ghost
        // below included) or [0, T-1] for a whole-lattice window
        const int zlo = ns > 1 ? -1 : 0;
        const int first = zlo + (parity_of((int)z0 + zlo, L) == fam ? 0 : 1);
        const uint32_t npl = first <= (int)T - 1 ? (uint32_t)((int)T - 1 - first) / 2 + 1 : 0u;
        const PhiloxFixed pfe = phase_draws(seed, kStEta, step, 0);
        if (FINE && fine_fits(4 * npl * Qc, dpc * Qc + duc)) {
          // fine: lane li takes cube li mod 4 of item li div 4
          const uint32_t c = li & 3u;
          const Walk it = walk_start(li >> 2, Qc, F(mQQ));
          if (it.pi < npl) {
            const uint32_t jj = fdiv(it.u, dQ), ii = it.u - jj * qL;
            const int zr = first + 2 * (int)it.pi;
            const uint32_t ox = 2 * ii + fam, oy = 2 * jj + fam;
            const int zg = (int)z0 + zr;
            const uint32_t zgw = (uint32_t)(zg < 0 ? zg + (int)L : zg);
            const uint32_t cube_ref = ii + hL * jj + (LL / 4) * zgw;
            const U4 r = draw_site(cube_ref, pfe);
            const uint32_t wb = (uint32_t)(zr + (int)gofs);
            const uint32_t wt = ns > 1 ? wb + 1 : (wb + 1 == T ? 0u : wb + 1);
            eta_cube_one_q(lds, lbase + wb * LL, lbase + wt * LL, thre, etab, hL, ox, oy, c, pick4(r, c));
          }
        } else {
          for (Walk it = walk_start(li, Qc, F(mQQ)); it.pi < npl; advance(it, Qc, dpc, duc)) {
            const uint32_t jj = fdiv(it.u, dQ), ii = it.u - jj * qL;
            const int zr = first + 2 * (int)it.pi;
            const uint32_t ox = 2 * ii + fam, oy = 2 * jj + fam;
            const int zg = (int)z0 + zr;
            const uint32_t zgw = (uint32_t)(zg < 0 ? zg + (int)L : zg);
            // cube index x/2 + (L/2)(y/2) + (L^2/4) z (gpu_water3D.cu:3405)
            const uint32_t cube_ref = ii + hL * jj + (LL / 4) * zgw;
            const U4 r = draw_site(cube_ref, pfe);
            const uint32_t wb = (uint32_t)(zr + (int)gofs);
            const uint32_t wt = ns > 1 ? wb + 1 : (wb + 1 == T ? 0u : wb + 1);
            if (fam)
              eta_cube_item_q<true>(lds, lbase + wb * LL, lbase + wt * LL, thre, etab, hL, ox, oy, r);
            else
              eta_cube_item_q<false>(lds, lbase + wb * LL, lbase + wt * LL, thre, etab, hL, ox, oy, r);
          }
        }
      }
      __syncthreads();
    }
    // ----------------------------------------------------- 12 half-sweeps
    for (uint32_t k = 0; k < 12; ++k) {
      const uint32_t sk = has ? uni(sch_s[gs][k]) : 0u;
      const bool wr = (sk >> 8) & 1u;
      const int arm = (int)(sk & 7u);
      const uint32_t q = ((sk >> 3) & 1u) ? 0u : 1u;
      const int ax = arm >> 1;
      if (run) {
        const uint32_t L = F(L), z0 = F(z0), T = F(T), ns = F(ns), hL = L / 2, qL = L / 4, LL = L * L;
        const uint64_t seed = ((uint64_t)F(seed_hi) << 32) | F(seed_lo);
        const uint32_t lbase = F(lds), gofs = ns > 1 ? 1u : 0u;
        const FastDiv dH{hL, F(mH)}, dQ{qL, F(mQ)};
        const uint32_t* thr = thr_s[gs];
        if (ax == 0) {
          sweep_inplane<0, FINE>(lds, li, L, T, z0, lbase + gofs * LL, seed, step, k, arm, q, dH, dQ, F(mP8), F(dp8),
                           F(du8), thr);
        } else if (ax == 1) {
          sweep_inplane<1, FINE>(lds, li, L, T, z0, lbase + gofs * LL, seed, step, k, arm, q, dH, dQ, F(mP8), F(dp8),
                           F(du8), thr);
        } else {
          // planes of the kept parity; the boundary plane whose neighbour
          // lies across the slab edge (top for +z, bottom for -z) is walked
          // last, as plane nint of the same walk
          const uint32_t Q4 = LL / 4, dp4 = F(dp4), du4 = F(du4);
          const uint32_t first = parity_of((int)z0, L) == q ? 0u : 1u;
          const uint32_t npl = first < T ? (T - first + 1) / 2 : 0u;
          const uint32_t zb = arm == 4 ? T - 1 : 0u;
          const bool bactive = npl > 0 && ((zb - first) & 1u) == 0;
          const int dz = arm == 4 ? (int)LL : -(int)LL;
          const uint32_t nint = bactive ? npl - 1 : npl;
          const uint32_t zi0 = first + (bactive && zb == 0 ? 2u : 0u);
          // the neighbour's plane as stored after its rd-th z half-sweep of
          // this step (rd = 0: the ghost, or the wrapped plane of a
          // whole-lattice window): tagged words of ring slot rd - 1; the
          // boundary plane's new words go to this slab's ring slot k (wr),
          // tagged
          const uint32_t rd = (sk >> 4) & 0xFu, tag = end_tag(step);
          const uint32_t stride = 64 * F(nw);
          const uint64_t* RG =
              rd ? reinterpret_cast<const uint64_t*>(
                       arm == 4 ? rings + F(ring_hi) + (half + ring_slot(rd - 1)) * LL
                                : rings + F(ring_lo) + (2 * kRingHalf + half + ring_slot(rd - 1)) * LL)
                 : nullptr;
          uint64_t* RW = bactive && wr ? reinterpret_cast<uint64_t*>(rings + F(ring) +
                                                                     ((arm == 4 ? 2 * kRingHalf : 0u) + half +
                                                                      ring_slot(k)) * LL)
                                       : nullptr;
          const int dzb = ns > 1 ? dz : (arm == 4 ? -(int)((T - 1) * LL) : (int)((T - 1) * LL));
          const uint64_t tw = (uint64_t)tag << 32;
          const PhiloxFixed pfz = phase_draws(seed, kStSigma, step, k);
          // item u of a plane is the quad with origin (u mod L/2, u div L/2),
          // word 4u of the plane
          if (FINE && 4 * npl * Q4 > stride && fine_fits(2 * npl * Q4, stride)) {
            // pairs: cells h0 = 2 (li mod 2) and h0 + 1 of item li div 2 (the
            // facing cells are the same pair of the plane above or below),
            // the boundary plane as plane nint
            const uint32_t h0 = 2 * (li & 1u);
            const Walk it = walk_start(li >> 1, Q4, F(mP4));
            bool ok = true;
            if (it.pi < npl) {
              const bool bnd = bactive && it.pi == nint;
              const uint32_t zr = bnd ? zb : zi0 + 2 * it.pi;
              const uint32_t y = fdiv(it.u, dH), x = it.u - y * hL;
              const uint32_t o = 4 * it.u + h0, cw = lbase + (zr + gofs) * LL + o;
              const uint2 w = *reinterpret_cast<const uint2*>(lds + cw);
              uint32_t n0, n1;
              if (bnd && RG) {
                const uint64_t* const src[2] = {RG + o, RG + o + 1};
                uint64_t v[2] = {ring_load64(src[0]), ring_load64(src[1])};
                ok = wait_tags(src, tag, v, status);
                n0 = (uint32_t)v[0];
                n1 = (uint32_t)v[1];
              } else {
                const uint2 n = *reinterpret_cast<const uint2*>(lds + cw + (uint32_t)(bnd ? dzb : dz));
                n0 = n.x;
                n1 = n.y;
              }
              const U4 rq = draw_site(x + L * y + LL * (z0 + zr), pfz);
              const uint32_t a = sigma_cell(w.x, n0, thr, arm, h0 ? rq.z : rq.x);
              const uint32_t b = sigma_cell(w.y, n1, thr, arm, h0 ? rq.w : rq.y);
              *reinterpret_cast<uint2*>(lds + cw) = make_uint2(a, b);
              if (bnd && RW) {
                ring_store64(RW + o, tw | a);
                ring_store64(RW + o + 1, tw | b);
              }
            }
            if (__ballot(!ok)) abort_s = 1;
          } else if (FINE && fine_fits(4 * npl * Q4, stride)) {
            // fine: one cell per lane (cell li mod 4 of item li div 4), the
            // boundary plane as plane nint
            const uint32_t h = li & 3u;
            const Walk it = walk_start(li >> 2, Q4, F(mP4));
            bool ok = true;
            if (it.pi < npl) {
              const bool bnd = bactive && it.pi == nint;
              const uint32_t zr = bnd ? zb : zi0 + 2 * it.pi;
              const uint32_t y = fdiv(it.u, dH), x = it.u - y * hL;
              const uint32_t o = 4 * it.u + h, cw = lbase + (zr + gofs) * LL + o;
              const uint32_t w = lds[cw];
              uint32_t wn;
              if (bnd && RG) {
                const uint64_t* const src[1] = {RG + o};
                uint64_t v[1] = {ring_load64(src[0])};
                ok = wait_tags(src, tag, v, status);
                wn = (uint32_t)v[0];
              } else {
                wn = lds[cw + (uint32_t)(bnd ? dzb : dz)];
              }
              const U4 rq = draw_site(x + L * y + LL * (z0 + zr), pfz);
              const uint32_t nwd = sigma_cell(w, wn, thr, arm, pick4(rq, h));
              lds[cw] = nwd;
              if (bnd && RW) ring_store64(RW + o, tw | nwd);
            }
            if (__ballot(!ok)) abort_s = 1;
          } else {
            // the lane's first boundary quad's four ring words are loaded
            // before the interior planes, so the ring's latency overlaps
            // them, and checked after
            const uint32_t bit = nint * Q4;
            const uint32_t fb = li >= bit ? li : li + fdiv(bit - li + stride - 1, FastDiv{stride, F(mS)}) * stride;
            const bool hasb = bactive && fb < bit + Q4;
            uint64_t pre[4] = {0, 0, 0, 0};
            if (RG && hasb) {
              const uint32_t o4 = 4 * (fb - bit);
#pragma unroll
              for (int m = 0; m < 4; ++m) pre[m] = ring_load64(RG + o4 + m);
            }
            Walk it = walk_start(li, Q4, F(mP4));
            for (; it.pi < nint; advance(it, Q4, dp4, du4)) {
              const uint32_t y = fdiv(it.u, dH), x = it.u - y * hL;
              const uint32_t zr = zi0 + 2 * it.pi;
              const uint32_t cw = lbase + (zr + gofs) * LL + 4 * it.u;
              uint32_t w[4], wn[4];
              quad_load(lds, cw, w);
              quad_load(lds, cw + (uint32_t)dz, wn);
              const U4 rq = draw_site(x + L * y + LL * (z0 + zr), pfz);
              sigma_quad_regs(w, wn, thr, arm, rq);
              quad_store(lds, cw, w);
            }
            if (bactive) {
              // the boundary plane, as plane nint of the same walk
              const uint32_t pl = lbase + (zb + gofs) * LL;
              bool ok = true;
              for (bool pf = true; it.pi == nint; advance(it, Q4, dp4, du4), pf = false) {
                const uint32_t y = fdiv(it.u, dH), x = it.u - y * hL;
                const uint32_t o4 = 4 * it.u;  // the quad's first word in the plane (and in a ring slot)
                uint32_t w[4], wn[4];
                quad_load(lds, pl + o4, w);
                if (!RG) {
                  quad_load(lds, pl + o4 + (uint32_t)dzb, wn);
                } else {
                  uint64_t v[4];
                  const uint64_t* const src[4] = {RG + o4, RG + o4 + 1, RG + o4 + 2, RG + o4 + 3};
#pragma unroll
                  for (int m = 0; m < 4; ++m) v[m] = pf ? pre[m] : ring_load64(src[m]);
                  ok = wait_tags(src, tag, v, status) && ok;
#pragma unroll
                  for (int m = 0; m < 4; ++m) wn[m] = (uint32_t)v[m];
                }
                const U4 rq = draw_site(x + L * y + LL * (z0 + zb), pfz);
                sigma_quad_regs(w, wn, thr, arm, rq);
                quad_store(lds, pl + o4, w);
                if (RW) {
#pragma unroll
                  for (int m = 0; m < 4; ++m) ring_store64(RW + o4 + m, tw | w[m]);
                }
              }
              if (__ballot(!ok)) abort_s = 1;
            }
          }
        }
      }
      __syncthreads();
    }
    // ------------------------------------- boundary planes at step end
    // (tagged words; the readers need no version.  The last step's end
    // also publishes its version, which the neighbours await before they
    // write back.)
    const bool last = si + 1 == nsteps;
    if (has && F(ns) > 1 && !abort_s) {
      const uint32_t L = F(L), LL = L * L, T = F(T);
      const uint64_t tag = (uint64_t)end_tag(step) << 32;
      uint64_t* dn = reinterpret_cast<uint64_t*>(rings + F(ring) + (half + kEndSlot) * LL);
      uint64_t* up = reinterpret_cast<uint64_t*>(rings + F(ring) + (2 * kRingHalf + half + kEndSlot) * LL);
      const uint32_t* bot = lds + F(lds) + LL;
      const uint32_t* top = lds + F(lds) + T * LL;
      // a one-plane slab's bottom and top plane are one plane: stored once,
      // in the lower half, where the slab above reads it too
      const uint32_t nst = T == 1 ? LL : 2 * LL;
      for (uint32_t i = li; i < nst; i += 64 * F(nw)) {
        if (i < LL)
          ring_store64(dn + i, tag | bot[i]);
        else
          ring_store64(up + (i - LL), tag | top[i - LL]);
      }
      if (last) ring_stores_done();
    }
    __syncthreads();
    if (last && has && F(ns) > 1 && wave == F(w0) && lane == 0 && !abort_s) {
      publish(ver + 2 * (G.first + gs), end_version(step));
      publish(ver + 2 * (G.first + gs) + 1, end_version(step));
    }
  }
  // the neighbours have read their windows (they published their last
  // step's version long after staging) before this workgroup overwrites its
  // planes in `words`
  if (has && F(ns) > 1 && nsteps > 0 && lane == 0 && wave == F(w0)) {
    const uint64_t want = end_version(step0 + nsteps - 1);
    if (!wait_version(ver + 2 * F(above), want, status) || !wait_version(ver + 2 * F(below) + 1, want, status))
      abort_s = 1;
  }
  __syncthreads();
  if (abort_s || !has) return;  // after the last barrier
  {
    // the slab's planes back to `words`, one quad per lane and item
    const uint32_t L = F(L), hL = L / 2, LL = L * L, Q4 = LL / 4, gofs = F(ns) > 1 ? 1u : 0u;
    uint32_t* Wp = words + F(woff) + F(z0) * LL;
    const uint4* Wl = reinterpret_cast<const uint4*>(lds + F(lds) + gofs * LL);
    const FastDiv dq{Q4, F(mP4)}, dH{hL, F(mH)};
    if ((hL & 3u) == 0) {
      // four quads along x per lane and item: 16-byte rows of `words`
      for (uint32_t i = 4 * li; i < F(T) * Q4; i += 256 * F(nw)) {
        const uint32_t p = fdiv(i, dq), r = i - p * Q4, y = fdiv(r, dH), x = r - y * hL;
        const uint4 q0 = Wl[i], q1 = Wl[i + 1], q2 = Wl[i + 2], q3 = Wl[i + 3];
        uint4* d = reinterpret_cast<uint4*>(Wp + p * LL + x + L * y);
        d[0] = make_uint4(q0.x, q1.x, q2.x, q3.x);
        d[hL * L / 4] = make_uint4(q0.y, q1.y, q2.y, q3.y);
        d[hL / 4] = make_uint4(q0.z, q1.z, q2.z, q3.z);
        d[(hL * L + hL) / 4] = make_uint4(q0.w, q1.w, q2.w, q3.w);
      }
    } else {
      for (uint32_t i = li; i < F(T) * Q4; i += 64 * F(nw)) {
        const uint32_t p = fdiv(i, dq), r = i - p * Q4, y = fdiv(r, dH), x = r - y * hL;
        const uint4 v = Wl[i];
        uint32_t* d = Wp + p * LL + x + L * y;
        d[0] = v.x;
        d[hL * L] = v.y;
        d[hL] = v.z;
        d[hL * L + hL] = v.w;
      }
    }
  }
#undef F
}

uint32_t resident_lds_limit() { return 160 * 1024 - kResidentStatic; }

static void set_resident_attribute() {
  // the dynamic-LDS limit is a per-device function attribute
  static std::atomic<uint64_t> attr_set{0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t bit = 1ull << (dev & 63);
  if (!(attr_set.load(std::memory_order_acquire) & bit)) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_resident<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)resident_lds_limit());
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_resident<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)resident_lds_limit());
    attr_set.fetch_or(bit, std::memory_order_acq_rel);
  }
}

hipError_t resident_occupancy(int* per_cu, uint32_t block, uint32_t lds_bytes, bool fine) {
  set_resident_attribute();
  return fine ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_resident<true>, (int)block, (size_t)lds_bytes)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_resident<false>, (int)block, (size_t)lds_bytes);
}

void launch_resident(hipStream_t st, const RGroupEnt* grp, uint32_t ngroups, uint32_t block, uint32_t lds_bytes,
                     const RSlab* slabs, const PointDev* pts, uint32_t* words, uint32_t* rings, uint64_t* ver,
                     const uint64_t* stepp, uint32_t step_off, uint32_t nsteps, uint32_t* status, bool fine) {
  set_resident_attribute();
  hipEvent_t t0, t1;
  get_launch_timing(&t0, &t1);
  auto* k = fine ? k_resident<true> : k_resident<false>;
  if (t0)
    hipExtLaunchKernelGGL(k, dim3(ngroups), dim3(block), lds_bytes, st, t0, t1, 0, grp, slabs, pts, words, rings, ver,
                          stepp, step_off, nsteps, status);
  else
    hipLaunchKernelGGL(k, dim3(ngroups), dim3(block), lds_bytes, st, grp, slabs, pts, words, rings, ver, stepp,
                       step_off, nsteps, status);
}

}  // namespace cvf
