// Device helpers shared by the kernel sources: the cube-move table and
// move, the sigma acceptance-table index and the in-plane quad geometry.
#pragma once

#include "cvf_core.h"

namespace cvf {

// Cube-move table: entry s holds the active-edge mask of solution s in bits
// 0-11 and, in bits 12 + 6b .. 17 + 6b, the eta bits (arms 0..5) that the
// mask gives corner b (b = bx + 2 by + 4 bz) on the arms it owns.
__device__ __forceinline__ uint64_t eta_table_entry(uint32_t s) {
  const uint32_t M = sol_mask(s);
  uint64_t t = M;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
    const uint32_t bits = (((M >> corner_edge(b, 0)) & 1u) << bx) | (((M >> corner_edge(b, 1)) & 1u) << (2 + by)) |
                          (((M >> corner_edge(b, 2)) & 1u) << (4 + bz));
    t |= (uint64_t)bits << (12 + 6 * b);
  }
  return t;
}

// One cube move on its 8 corner words w (corner b = bx + 2 by + 4 bz):
// proposal floor(9u), Delta N_HB over the 12 edges, accept with
// min(1, exp(beta Jeff DeltaN)); store(b, word) writes a changed corner.
template <class Store>
__device__ __forceinline__ void eta_cube_move(const uint32_t (&w)[8], const uint64_t* thr, const uint64_t* etab,
                                              uint32_t rw, Store&& store) {
  uint32_t D = 0, E = 0;
#pragma unroll
  for (int jj = 0; jj < kEdges; ++jj) {
    const int A = edge_start(jj), ax = edge_axis(jj), B = A | (1 << ax);
    D |= (uint32_t)(sigma_of(w[A], 2 * ax) == sigma_of(w[B], 2 * ax + 1)) << jj;
    E |= eta_of(w[A], 2 * ax) << jj;
  }
  uint32_t prop, acc;
  split_word(rw, 9, prop, acc);
  const uint64_t ent = etab[prop];
  const uint32_t M = (uint32_t)ent & 0xFFFu;
  const int dn = popc32(M & D) - popc32(E & D);
  if ((uint64_t)acc < thr[dn + 4] && M != E) {
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t own = (1u << (b & 1)) | (1u << (2 + ((b >> 1) & 1))) | (1u << (4 + ((b >> 2) & 1)));
      const uint32_t bits = (uint32_t)(ent >> (12 + 6 * b)) & 0x3Fu;
      store(b, (w[b] & ~(own << kEtaShift)) | (bits << kEtaShift));
    }
  }
}

// Cube moves on the four cubes of a cube item: origins (ox + a L/2, oy +
// b L/2) on bottom plane B and top plane U (L x L words each, periodic in
// x and y), random words r.x .. r.w in the order (0,0), (0,1), (1,0),
// (1,1).  The four cubes are disjoint, so all 32 corner words are loaded
// before the first move.
__device__ __forceinline__ void eta_cube_item(uint32_t* lds, uint32_t B, uint32_t U, const uint64_t* thr,
                                              const uint64_t* etab, uint32_t L, uint32_t ox, uint32_t oy,
                                              const U4& r) {
  const uint32_t hL = L / 2;
  uint32_t idx[4][8], w[4][8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t cx = ox + ((c & 2) ? hL : 0u), cy = oy + ((c & 1) ? hL : 0u);
    const uint32_t x1 = cx + 1 == L ? 0 : cx + 1;
    const uint32_t y1 = cy + 1 == L ? 0 : cy + 1;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t x = (b & 1) ? x1 : cx, y = (b & 2) ? y1 : cy;
      idx[c][b] = ((b & 4) ? U : B) + x + L * y;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int b = 0; b < 8; ++b) w[c][b] = lds[idx[c][b]];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t rw = c == 0 ? r.x : c == 1 ? r.y : c == 2 ? r.z : r.w;
    eta_cube_move(w[c], thr, etab, rw, [&](int b, uint32_t v) { lds[idx[c][b]] = v; });
  }
}

// Sigma acceptance thresholds as 32-bit values min(thr, 2^32 - 1).  Exact:
// the acceptance word of the q = 6 split is (6w) mod 2^32, always even, so
// it is below 2^32 - 1 exactly when it is below 2^32.
__device__ __forceinline__ uint32_t sigma_thr32(uint64_t t) {
  return t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
}

// Acceptance-table index (h + 1) * 11 + d + 5 of a chessboard sigma move of
// arm `arm` from sa to sp (sp != sa; any in-range index otherwise) on word w
// with facing neighbour word wn, where h = eta ? delta(sp, snb) -
// delta(sa, snb) : 0 and d = count_eq(w, sp) - count_eq(w, sa) + 1.
// Branch-free: with sp != sa at most one of the deltas is 1, so (h + 1) * 11
// is a two-way select; d + 5 = #(fields != sa) + #(fields == sp), each from
// the zero-field test ((y & 011) + 011 | y) & 100 on every 3-bit field of
// y = w ^ (s * 001) (no carry leaves a field), the counts chained through
// the popcount's addend.
// popcount(x) + acc in one v_bcnt_u32_b32 (written out so that the compiler
// does not spread the table's x4 scaling over the terms of the sum)
__device__ __forceinline__ uint32_t popc_add(uint32_t x, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
  return r;
}

__device__ __forceinline__ uint32_t sigma_thr_index(uint32_t w, uint32_t wn, int arm, uint32_t sa, uint32_t sp) {
  constexpr uint32_t kLo2 = 0x1B6DBu, kHi = 0x24924u;  // bits 0-1, bit 2 of each field
  const uint32_t snb = sigma_of(wn, arm ^ 1);
  const bool eta = (w >> (kEtaShift + arm)) & 1u;
  const uint32_t h11 = eta && sp == snb ? 22u : eta && sa == snb ? 0u : 11u;
  const uint32_t ya = w ^ (sa * kFieldLsb), yp = w ^ (sp * kFieldLsb);
  const uint32_t ne_a = (((ya & kLo2) + kLo2) | ya) & kHi;
  const uint32_t eq_p = ~(((yp & kLo2) + kLo2) | yp) & kHi;
  return popc_add(ne_a, popc_add(eq_p, h11));
}

// Cells and facing neighbours (window indices) of the in-plane (x or y arm)
// quad with origin (x, y) on the plane starting at word `base`: cells (x, y),
// (x, y + L/2), (x + L/2, y), (x + L/2, y + L/2) with x, y < L/2.  Along the
// arm's axis only the coordinate at 0 (-1 step) or L-1 (+1 step) wraps, i.e.
// the low cells when x (y) = 0 and the high cells when x (y) = L/2 - 1.
__device__ __forceinline__ void quad_plane_cells(uint32_t base, uint32_t L, uint32_t x, uint32_t y, int arm,
                                                 int ax, uint32_t* cell, uint32_t* nb) {
  const int iL = (int)L, hL = iL / 2, cq = ax == 0 ? (int)x : (int)y, st = ax == 0 ? 1 : iL;
  const int wrap = (iL - 1) * st;
  int off_lo, off_hi;
  if (arm & 1) {
    off_lo = cq == 0 ? wrap : -st;
    off_hi = -st;
  } else {
    off_lo = st;
    off_hi = cq == hL - 1 ? -wrap : st;
  }
  const uint32_t b0 = base + x + L * y, dy = (uint32_t)hL * L, dx = (uint32_t)hL;
  cell[0] = b0;
  cell[1] = b0 + dy;
  cell[2] = b0 + dx;
  cell[3] = b0 + dx + dy;
  // the cells on the high side of the arm's axis: 2, 3 for x, 1, 3 for y
  nb[0] = cell[0] + off_lo;
  nb[1] = cell[1] + (ax == 0 ? off_lo : off_hi);
  nb[2] = cell[2] + (ax == 0 ? off_hi : off_lo);
  nb[3] = cell[3] + off_hi;
}

}  // namespace cvf
