// Device-side data structures and the host launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvf_core.h"

namespace cvf {

constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

// Immutable per-state-point description (device copy).
struct PointDev {
  uint64_t seed;
  uint32_t L, N, hL;
  uint32_t woff;       // offset of the point's words in the state array
  uint32_t loff;       // offset of the point's 6N labels
  uint32_t shell_off;  // offset of its shell table
  int32_t jeff_sign;   // sign of Jeff (SW covalent rule, gpu_water3D.cu:2399-2424)
  uint32_t pad;
  VolParams vp;
  uint64_t thr_sigma[33];  // index (eta*dhb + 1)*11 + dsig + 5
  uint64_t thr_eta[9];     // index DeltaN_HB + 4
  uint64_t thr_sw[2];      // covalent, intra-molecule link activation
};

// Mutable per-state-point device state.
struct PointState {
  uint64_t step;  // next MC step index of this point
  VolState vol;
};

// One workgroup's work: the state point, the first work item, and every
// per-point scalar the kernels need, so that one scalar load of the entry
// starts a workgroup (96 bytes).
struct BlockEnt {
  int32_t point;  // -1: padding entry
  uint32_t start;  // first work item (or first plane of a slab)
  uint32_t L, N, woff, loff;
  uint32_t seed_lo, seed_hi;
  uint32_t mL, mH, mQ;  // fast-division multipliers for L, L/2, L/4
  int32_t jeff_sign;
  uint32_t thb_lo, thb_hi, ts_lo, ts_hi;  // SW link thresholds
  uint32_t mLL, mP8, mP4, mC;  // multipliers for L^2, L^2/8, L^2/4, L^2/16
  uint32_t T;                  // slab thickness (planes) of the fused kernel
  uint32_t flags;              // kEntNoChess: the point runs without the chessboard
  uint32_t pad[2];
};

constexpr uint32_t kEntNoChess = 1;

constexpr int kChunk = 6;      // half-sweeps per fused kernel (default)
constexpr int kMaxChunk = 12;  // half-sweeps per fused kernel at most: a whole Metropolis pass
constexpr uint32_t kObserveSpan = 4096;  // cells per k_observe workgroup (256 threads)
constexpr int kMaxSlab = 16;  // planes per slab of the fused kernel

struct Launch {
  const BlockEnt* tab;
  uint32_t nblocks;      // padded to a multiple of kNumXcd
  uint32_t block;        // threads per workgroup
  const uint64_t* step;  // the group's step counter (first step of the captured sequence)
  uint32_t step_off;     // this launch's step within the sequence
};

// Per-kernel timing for profiling: while set, launches go through
// hipExtLaunchKernelGGL with these start/stop events (nullptr to clear).
void set_launch_timing(hipEvent_t start, hipEvent_t stop);

// Launchers (all asynchronous on `st`).
// after nsteps lattice steps of a group: *gstep += nsteps and, with `moves`,
// the device volume moves of those steps for the listed points
void launch_vol_chain(hipStream_t st, const PointDev* pts, const int32_t* plist, int npts, PointState* ps,
                      uint64_t* gstep, uint32_t nsteps, const float* dist, const int32_t* freq,
                      int precision, int moves);
void launch_eta(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, int family);
void launch_sigma(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, uint32_t sub);
// one colour pass of the fast schedule (pass = 2 x Metropolis iteration +
// colour), over the half-sweep table (N/8 quads per point)
void launch_sigma_f(hipStream_t st, Launch ln, const PointDev* pts, uint32_t* words, uint32_t pass);
// Fused z-slab kernel: optional eta sweep + half-sweeps k0 .. k0+nk-1 of the
// step on LDS-staged slabs, reading `src` and writing `dst`; any_nc: the
// table holds non-chessboard slabs.
void launch_chunk(hipStream_t st, Launch ln, const PointDev* pts, const uint32_t* src, uint32_t* dst,
                  uint32_t k0, uint32_t nk, int with_eta, uint32_t lds_bytes, bool any_nc);
// largest dynamic LDS a fused launch may use
uint32_t chunk_lds_limit();
void launch_zero_u32(hipStream_t st, uint32_t* p, size_t n);
void launch_zero3_u32(hipStream_t st, uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc);
void launch_observe(hipStream_t st, Launch ln, const uint32_t* words, unsigned long long* counters);
void launch_sw_local(hipStream_t st, Launch ln, const uint32_t* words, uint32_t* parent, uint8_t* cov,
                     uint32_t jc);
void launch_sw_union(hipStream_t st, Launch ln, uint32_t* parent, const uint8_t* cov);
void launch_sw_flip(hipStream_t st, Launch ln, uint32_t* words, uint32_t* parent, uint32_t jc);
// sigma histogram as kHistReplicas partial histograms of 6 counters
// (kHistWords counters, zeroed by the caller) that the caller sums
constexpr int kHistReplicas = 8;
constexpr int kHistWords = 6 * kHistReplicas;
void launch_sigma_hist(hipStream_t st, const uint32_t* W, uint32_t N, unsigned long long* counts);
// the same into slot (*gstep + step_off - *base) of a per-step series of
// `span` histograms (graph-replayable: the slot is read on the device)
void launch_sigma_hist_slot(hipStream_t st, const uint32_t* W, uint32_t N, unsigned long long* series,
                            const uint64_t* gstep, uint32_t step_off, const uint64_t* base, uint32_t span);
// cluster statistics of one point (labels at `label`, 6N nodes)
void launch_cluster_sizes(hipStream_t st, const uint32_t* label, uint32_t nn, int32_t* size);
// Cluster sizes size[0..nn) (zero: no cluster): the count into *total and
// each workgroup's offset into bcount (compact_blocks(nn) words), and the
// sizes' moments bmom[4 nb .. 4 nb + 3] = sum, sum of squares, largest |
// second largest << 32 (bmom: 4 (compact_blocks(nn) + 1) words); then, with
// those offsets, the non-zero sizes in label order into out[0..*total)
constexpr uint32_t kCompactSpan = 1024;
uint32_t compact_blocks(uint32_t nn);
void launch_compact_count(hipStream_t st, const int32_t* size, uint32_t nn, uint32_t* bcount, uint32_t* total,
                          uint64_t* bmom);
void launch_compact_write(hipStream_t st, const int32_t* size, uint32_t nn, const uint32_t* bcount, int32_t* out);
void launch_perc_slots(hipStream_t st, const uint32_t* label, uint32_t nn, const int32_t* size,
                       int32_t minsize, int32_t* slot, uint32_t* nslot, uint32_t max_slots);
void launch_perc_cover(hipStream_t st, const uint32_t* label, uint32_t nn, uint32_t L,
                       const int32_t* size, int32_t minsize, const int32_t* slot, uint32_t* cover,
                       uint32_t max_slots);
void launch_perc_count(hipStream_t st, const uint32_t* cover, const uint32_t* nslot, uint32_t L,
                       uint32_t max_slots, uint32_t* nperc);
void launch_vdw_eval(hipStream_t st, const float* dist, const int32_t* freq, int nshell,
                     float r_cutoff, int N, const double* r, int n_r, int precision, double* out);
// vdW pair form (validation of the shell table): per-workgroup partial sums
// over ordered pairs, vdw_pair_blocks(L) per radius, radius-major in `part`
uint32_t vdw_pair_blocks(int L);
void launch_vdw_pairs(hipStream_t st, int L, float r_cutoff, const double* r, int n_r, int precision, double* part);
// the events set by set_launch_timing (nullptr when launches are not timed)
void get_launch_timing(hipEvent_t* start, hipEvent_t* stop);

// ------------------------------------------------- resident Metropolis kernel
// One z-slab of one state point held in a resident workgroup's LDS
// (cvf_resident.hip), worked on by its own waves of the workgroup.  Its
// window: the ghost of the plane below, its T planes, the ghost of the
// plane above (no ghosts when ns = 1: the whole lattice, periodic in LDS).  Its rings in `rings` (words): [ring, ring +
// 2 kRingHalf L^2) its bottom plane as published for the slab below, then
// its top plane for the slab above, each a step-parity half of kRingHalf
// L^2-word slots holding 64-bit tagged words (two slots per plane);
// versions ver[2 i] (bottom), ver[2 i + 1] (top) for slab i.
struct RSlab {
  int32_t point;
  uint32_t L, z0, T, ns;      // ns: slabs of the point
  uint32_t woff, lds;         // the point's words; the window's first LDS word
  uint32_t seed_lo, seed_hi;
  uint32_t mH, mQ, mP8, mP4, mQQ;  // fast-division multipliers of L/2, L/4, L^2/8, L^2/4, (L/4)^2
  int32_t above, below;       // neighbouring slabs (ns > 1), indices into the slab array
  uint32_t ring;              // this slab's rings
  uint32_t ring_hi, ring_lo;  // the rings of `above` and `below`
  uint32_t w0, nw;            // its waves [w0, w0 + nw) of the workgroup
  // a lane's walk over a phase's items advances by 64 nw items: dp planes
  // and du items of L^2/8 (in-plane quads), L^2/4 (z quads) or (L/4)^2
  // (four-cube items) per plane
  uint32_t dp8, du8, dp4, du4, dpc, duc;
  uint32_t mS;  // fast-division multiplier of that stride, 64 nw
};
static_assert(sizeof(RSlab) % 16 == 0, "RSlab is copied in words");
// Workgroup g of a resident launch takes slabs [first, first + count).
struct RGroupEnt {
  uint32_t first, count;
};
constexpr uint32_t kMaxRSlabs = 8;   // slabs per resident workgroup
constexpr uint32_t kRingHalf = 32;   // ring slots per step parity: 2k, 2k + 1 after half-sweep k; 30, 31 at step end
constexpr uint32_t kResidentStatic = 6144;  // static LDS reserved for a resident workgroup
// largest dynamic LDS of a resident workgroup (its static arrays take the
// rest of the CU's 160 KB; each further workgroup on a CU takes another
// kResidentStatic)
uint32_t resident_lds_limit();
// resident workgroups of `block` threads and `lds_bytes` of dynamic LDS that
// fit on one CU at once
hipError_t resident_occupancy(int* per_cu, uint32_t block, uint32_t lds_bytes, bool fine);
// fine: the instantiation for plans of thin slabs (at most two planes):
// in-plane half-sweeps whose quads fit half of a slab's lanes run two cells
// per lane (one, below 64 quads), cube families whose cubes fit the lanes one cube per lane, and
// z half-sweeps one cell per lane where that fits, else two where that
// does, instead of one quad (one four-cube item) per lane
void launch_resident(hipStream_t st, const RGroupEnt* grp, uint32_t ngroups, uint32_t block, uint32_t lds_bytes,
                     const RSlab* slabs, const PointDev* pts, uint32_t* words, uint32_t* rings, uint64_t* ver,
                     const uint64_t* stepp, uint32_t step_off, uint32_t nsteps, uint32_t* status, bool fine);

}  // namespace cvf
