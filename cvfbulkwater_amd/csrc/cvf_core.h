// Core definitions shared by the HIP kernels and the host side of the engine
// (everything here is __host__ __device__ and free of dynamic allocation).
//
// Lattice state: one 32-bit word per cell, cell index x + L*y + L*L*z
// (gpu_water3D.cu:821-840 uses the same ordering for its arm-major arrays):
//   bits  0..17  sigma_a, 3 bits per arm a = 0..5 (0:+x 1:-x 2:+y 3:-y 4:+z 5:-z)
//   bits 18..23  eta_a, one bit per arm
//   bits 24..31  zero
// The eta bits always equal the solution of the 2x2x2 cube that owns the
// bond (chessboard.h:211-262), so no cube-state array is kept.
//
// Random numbers: Philox4x32-10 (Salmon et al., SC'11), key = seed,
// counter = (site, step_lo, stream<<24 | sub, step_hi).  This replaces the
// reference's pre-generated cuRAND arrays (gpu_water3D.cu:584-625) and host
// MT19937 / xorshift1024* streams, and makes every accept/reject an integer
// comparison that the CPU oracle reproduces exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvf {

constexpr int kArms = 6;
constexpr int kEdges = 12;
constexpr int kNSol = 9;
constexpr uint32_t kSigmaBits = 0x3FFFFu;
constexpr uint32_t kEtaShift = 18;
// bits 24-31: per-cell HB counter of the non-chessboard mode (uint8 arithmetic)
constexpr uint32_t kNhbShift = 24;
constexpr uint32_t kLowBits = 0x00FFFFFFu;
constexpr uint32_t kFieldLsb = 0x9249u;  // bit 0 of each 3-bit sigma field
constexpr int kMaxL = 252;               // fast division exact for L <= 255

enum Stream : uint32_t {
  kStSched = 1,
  kStSigma = 2,
  kStEta = 3,
  kStVolume = 4,
  kStSwBond = 5,
  kStSwFlip = 6,
  kStInitSigma = 7,
  kStInitCube = 8,
  kStSigmaF = 9,  // colour passes of the fast schedule
};

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * b) >> 32);
}

// a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) on gfx950, where the
// compiler otherwise issues two v_xor_b32.
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// Rounds R0 .. 9 of Philox4x32-10 on (c0, c1, c2, c3), (k0, k1) being the
// key of round R0.
template <int R0>
__host__ __device__ __forceinline__ U4 philox_rounds(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                     uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = R0; r < 10; ++r) {
    if (r > R0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return U4{c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1) {
  return philox_rounds<0>(c0, c1, c2, c3, k0, k1);
}

// Philox4x32-10 for many counters that differ in word c0 only.  Neither
// round 0's product of c2 nor round 1's product of round 0's word 0 (a
// function of c1, c2 and the key) involves c0: philox_fixed computes them
// once, philox_site completes one counter from them.
// philox_site(c0, philox_fixed(c1, c2, c3, k0, k1)) == philox(c0, c1, c2, c3, k0, k1).
struct PhiloxFixed {
  uint32_t t0, t1, t2, t3;  // round 0: word 2's term c3 ^ k1; round 1: the terms of words 0, 2 and word 3
  uint32_t k0, k1;          // key of round 2
};
__host__ __device__ __forceinline__ PhiloxFixed philox_fixed(uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                             uint32_t k1) {
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
  const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;  // round 0, word 0
  const uint32_t k0b = k0 + kPhiloxW0, k1b = k1 + kPhiloxW1;
  const uint64_t q0 = (uint64_t)kPhiloxM0 * n0;        // round 1, word 0's product
  return PhiloxFixed{c3 ^ k1, (uint32_t)p1 ^ k0b, (uint32_t)(q0 >> 32) ^ k1b, (uint32_t)q0, k0b + kPhiloxW0,
                     k1b + kPhiloxW1};
}
__host__ __device__ __forceinline__ U4 philox_site(uint32_t c0, const PhiloxFixed& f, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;                  // round 0, word 2's product
  const uint32_t y2 = (uint32_t)(p0 >> 32) ^ f.t0, y3 = (uint32_t)p0;
  const uint64_t p1 = (uint64_t)kPhiloxM1 * y2;                  // round 1, word 0's product
  return philox_rounds<2>((uint32_t)(p1 >> 32) ^ f.t1, (uint32_t)p1, f.t2 ^ y3, f.t3, k0, k1);
}
__host__ __device__ __forceinline__ U4 philox_site(uint32_t c0, const PhiloxFixed& f) {
  return philox_site(c0, f, f.k0, f.k1);
}

// One draw of a stream: site, step, stream id, sub-index.
__host__ __device__ __forceinline__ U4 draw(uint64_t seed, uint32_t stream, uint32_t site,
                                            uint64_t step, uint32_t sub) {
  return philox(site, (uint32_t)step, (stream << 24) | sub, (uint32_t)(step >> 32),
                (uint32_t)seed, (uint32_t)(seed >> 32));
}

// floor(n * x / 2^32): integer form of (int)(n*u) for u in [0,1)
// (gpu_water3D.cu:1968-1969, 3411-3412 use (int)(n*u) with u in (0,1]).
__host__ __device__ __forceinline__ uint32_t uniform_int(uint32_t x, uint32_t n) {
  return mulhi32(x, n);
}

// One Philox word per site update: X = q*w (35-bit); the proposal is
// X >> 32 = floor(q u) and the acceptance uniform is X mod 2^32, the
// fractional part of q u in 32-bit fixed point (u = w / 2^32).
__host__ __device__ __forceinline__ void split_word(uint32_t w, uint32_t q, uint32_t& prop,
                                                   uint32_t& acc) {
  const uint64_t X = (uint64_t)w * q;
  prop = (uint32_t)(X >> 32);
  acc = (uint32_t)X;
}

// 53-bit uniform double in [0,1).
__host__ __device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// ----------------------------------------------------------- packed word
__host__ __device__ __forceinline__ uint32_t sigma_of(uint32_t w, int a) { return (w >> (3 * a)) & 7u; }
__host__ __device__ __forceinline__ uint32_t eta_of(uint32_t w, int a) {
  return (w >> (kEtaShift + a)) & 1u;
}
__host__ __device__ __forceinline__ int popc32(uint32_t v) {
#ifdef __HIP_DEVICE_COMPILE__
  return __popc(v);
#else
  return __builtin_popcount(v);
#endif
}
// number of arms whose sigma equals v
__host__ __device__ __forceinline__ int count_eq(uint32_t w, uint32_t v) {
  const uint32_t x = (w & kSigmaBits) ^ (v * kFieldLsb);
  const uint32_t nz = (x | (x >> 1) | (x >> 2)) & kFieldLsb;
  return 6 - popc32(nz);
}

// Fast exact division by a runtime constant d >= 1: floor(t/d) =
// mulhi(t, ceil(2^32/d)), exact whenever t*d < 2^32 (d = 1 is passed through).
struct FastDiv {
  uint32_t d, m;
};
__host__ inline FastDiv make_fastdiv(uint32_t d) {
  return FastDiv{d, d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + d - 1) / d)};
}
__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t t, FastDiv f) {
  return f.d <= 1 ? t : mulhi32(t, f.m);
}

// ------------------------------------------------- cube (chessboard) tables
// 12 edges of the 2x2x2 cube with origin v1 (gpu_water3D.cu:3136-3176);
// corner index b = bx + 2*by + 4*bz.  Edge j starts at corner kEdgeStart[j]
// and runs along axis kEdgeAxis[j].
// Active-edge masks of the 9 solutions (chessboard.h:211-262).
__host__ __device__ __forceinline__ uint32_t sol_mask(uint32_t s) {
  // each solution switches off 4 edges; mask = 0xFFF & ~off
  const uint32_t off0 = (1u << 2) | (1u << 3) | (1u << 6) | (1u << 9);
  const uint32_t off1 = (1u << 2) | (1u << 3) | (1u << 5) | (1u << 11);
  const uint32_t off2 = (1u << 0) | (1u << 6) | (1u << 8) | (1u << 9);
  const uint32_t off3 = (1u << 0) | (1u << 5) | (1u << 8) | (1u << 11);
  const uint32_t off4 = (1u << 1) | (1u << 3) | (1u << 7) | (1u << 11);
  const uint32_t off5 = (1u << 1) | (1u << 4) | (1u << 7) | (1u << 10);
  const uint32_t off6 = (1u << 2) | (1u << 4) | (1u << 5) | (1u << 10);
  const uint32_t off7 = (1u << 1) | (1u << 4) | (1u << 8) | (1u << 9);
  const uint32_t off8 = (1u << 0) | (1u << 6) | (1u << 7) | (1u << 10);
  uint32_t off = s == 0 ? off0 : s == 1 ? off1 : s == 2 ? off2 : s == 3 ? off3 : s == 4 ? off4
               : s == 5 ? off5 : s == 6 ? off6 : s == 7 ? off7 : off8;
  return 0xFFFu & ~off;
}
// edge start corner and axis
__host__ __device__ __forceinline__ int edge_start(int j) {
  // 0:v1 1:v1 2:v1 3:v2 4:v2 5:v3 6:v3 7:v4 8:v4 9:v6 10:v7 11:v8
  // v1=(0,0,0)=0 v2=(1,0,0)=1 v3=(0,1,0)=2 v4=(0,0,1)=4 v6=(0,1,1)=6 v7=(1,0,1)=5 v8=(1,1,0)=3
  const uint64_t tab = 0x356442211000ull;  // nibble j = start corner
  return (int)((tab >> (4 * j)) & 0xF);
}
__host__ __device__ __forceinline__ int edge_axis(int j) {
  // x-edges {0,6,8,9}, y-edges {1,4,7,10}, z-edges {2,3,5,11}
  return (j == 0 || j == 6 || j == 8 || j == 9) ? 0 : (j == 1 || j == 4 || j == 7 || j == 10) ? 1 : 2;
}
// edge owned by corner b along axis ax (the edge lies on the cube face that
// contains b; b is its start if b's ax-bit is 0, its end otherwise)
__host__ __device__ __forceinline__ int corner_edge(int b, int ax) {
  const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
  if (ax == 0) {
    const int t[4] = {0, 6, 8, 9};  // (by,bz)
    return t[by + 2 * bz];
  } else if (ax == 1) {
    const int t[4] = {1, 4, 7, 10};  // (bx,bz)
    return t[bx + 2 * bz];
  }
  const int t[4] = {2, 3, 5, 11};  // (bx,by)
  return t[bx + 2 * by];
}

// ------------------------------------------- deterministic exp / log (FP64)
// Written for this project so that the host (plain form) and the device
// (batched form) volume chains are bit-identical: only IEEE +,-,*,/ and
// exponent-field arithmetic are used, with contraction disabled.
#pragma clang fp contract(off)

__host__ __device__ inline double as_double(uint64_t u) {
  union {
    uint64_t u;
    double d;
  } v;
  v.u = u;
  return v.d;
}
__host__ __device__ inline uint64_t as_u64(double d) {
  union {
    uint64_t u;
    double d;
  } v;
  v.d = d;
  return v.u;
}
__host__ __device__ inline double pow2i(int k) {  // 2^k for -1022 <= k <= 1023
  return as_double((uint64_t)(k + 1023) << 52);
}
__host__ __device__ inline double scale2(double p, int k) {
  if (k > 1000) {
    p *= pow2i(1000);
    k -= 1000;
  } else if (k < -1000) {
    p *= pow2i(-1000);
    k += 1000;
  }
  return p * pow2i(k);
}

// exp(x): Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2 (ln2 split into a
// 32-bit head and a tail so that k*ln2_hi is exact), then the degree-13
// Taylor polynomial of e^r by Horner (truncation < 5e-18 relative), then
// scaling by 2^k.  Accuracy about 1 ulp.
__host__ __device__ inline double det_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return as_double(0x7FF0000000000000ull);
  if (x < -745.2) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double t = x * inv_ln2;
  const double kd = (double)(int64_t)(t < 0 ? t - 0.5 : t + 0.5);
  const double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return scale2(p, (int)kd);
}

// log(x) for finite x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// log(m) = 2 atanh(f), f = (m-1)/(m+1), |f| <= 0.1716, odd series to f^25.
__host__ __device__ inline double det_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -as_double(0x7FF0000000000000ull) : as_double(0x7FF8000000000000ull);
  if (x != x || x == as_double(0x7FF0000000000000ull)) return x;
  uint64_t b = as_u64(x);
  int e = (int)((b >> 52) & 0x7FF);
  if (e == 0) {  // subnormal
    x *= 18014398509481984.0;  // 2^54
    b = as_u64(x);
    e = (int)((b >> 52) & 0x7FF) - 54;
  }
  e -= 1023;
  double m = as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double s = f * f;
  double R = 1.0 / 25.0;
  R = R * s + 1.0 / 23.0;
  R = R * s + 1.0 / 21.0;
  R = R * s + 1.0 / 19.0;
  R = R * s + 1.0 / 17.0;
  R = R * s + 1.0 / 15.0;
  R = R * s + 1.0 / 13.0;
  R = R * s + 1.0 / 11.0;
  R = R * s + 1.0 / 9.0;
  R = R * s + 1.0 / 7.0;
  R = R * s + 1.0 / 5.0;
  R = R * s + 1.0 / 3.0;
  R = R * s;
  const double logm = 2.0 * f + 2.0 * f * R;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double ed = (double)e;
  return ed * ln2_hi + (ed * ln2_lo + logm);
}

// ------------------------------------------------------- van der Waals
// E_vdW(r) = sum_k f_k [(r d_k)^-12 - (r d_k)^-6 - shift], over shells with
// r*d_k < R_c (calculate_energy_vdW, gpu_water3D.cu:1784-1805), FP64 in the
// reference's operation order.
__host__ __device__ inline double vdw_energy_fp64(double r, float r_cutoff, const float* dist,
                                                  const int32_t* freq, int nshell) {
  double tmp = 1.0 / (r_cutoff);
  tmp = tmp * tmp * tmp;
  double att = tmp * tmp;
  double rep = att * att;
  const double shift = rep - att;
  double ret = 0;
  for (int i = 0; i < nshell; ++i)
    if (r * dist[i] < r_cutoff) {
      tmp = 1.0 / (r * dist[i]);
      tmp = tmp * tmp * tmp;
      att = tmp * tmp;
      rep = att * att;
      ret += freq[i] * (rep - att - shift);
    }
  return ret;
}

// Volume-chain state and the constants it needs.
struct VolState {
  double r_cell, V, E_vdw;
  uint32_t accepted, pad;
};

struct VolParams {
  uint64_t seed;
  int32_t L, N;
  float T, P, r_cutoff;
  int32_t nshell;
  double beta;
};

// Trial radius and the two uniforms of the volume move at `step`
// (gpu_water3D.cu:1740-1748): r' = r + 0.01*2*(u1 - 0.5), redrawn while
// r' < 0 (LJ_InfiniteBarrier = 0).
__host__ __device__ inline double volume_trial(const VolState& st, const VolParams& p, uint64_t step,
                                               double* u2) {
  double trial;
  uint32_t attempt = 0;
  U4 r;
  const double eps = 0.01;
  do {
    r = draw(p.seed, kStVolume, attempt, step, 0);
    trial = st.r_cell + eps * 2 * (u53(r.x, r.y) - 0.5);
    ++attempt;
  } while (trial < 0.0);
  *u2 = u53(r.z, r.w);
  return trial;
}

// NPT acceptance of gpu_water3D.cu:1750-1781 given the trial energy:
// dE = (E' - E) - N T ln(V'/V), accept iff u2 < exp(-beta (dE + P dV)).
__host__ __device__ inline bool volume_accept(VolState& st, const VolParams& p, double trial,
                                              double trial_energy, double u2) {
  const double L = (double)p.L;
  const double tv = trial * trial * trial * L * L * L;
  const double dV = tv - st.V;
  const float NT = (float)p.N * p.T;
  const double dE = (trial_energy - st.E_vdw) - (double)NT * det_log(tv / st.V);
  if (u2 < det_exp(-p.beta * (dE + (double)p.P * dV))) {
    st.r_cell = trial;
    st.V = tv;
    st.E_vdw = trial_energy;
    st.accepted += 1;
    return true;
  }
  return false;
}

#pragma clang fp contract(on)

}  // namespace cvf
