// CPU test probe: exposes the engine's shared __host__ __device__ core
// (cvf_core.h) and host helpers (cvf_host.cpp) through plain C functions so
// that the CPU test suite can check them without a GPU, plus rocRAND's
// Philox4x32-10 round function as an independent implementation.
#include <rocrand/rocrand_philox4x32_10.h>

#include <vector>

#include "cvf_core.h"
#include "cvf_host.h"

using namespace cvf;

namespace {
struct RocPhilox : rocrand_device::philox4x32_10_engine {
  uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
}  // namespace

extern "C" {

void probe_philox(const uint32_t* c, const uint32_t* k, uint32_t* out) {
  const U4 r = philox(c[0], c[1], c[2], c[3], k[0], k[1]);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

// the split form the kernels use for counters that differ in word 0 only
void probe_philox_site(const uint32_t* c, const uint32_t* k, uint32_t* out) {
  const U4 r = philox_site(c[0], philox_fixed(c[1], c[2], c[3], k[0], k[1]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void probe_rocrand_philox(const uint32_t* c, const uint32_t* k, uint32_t* out) {
  RocPhilox e;
  uint4 cc; cc.x = c[0]; cc.y = c[1]; cc.z = c[2]; cc.w = c[3];
  uint2 kk; kk.x = k[0]; kk.y = k[1];
  const uint4 r = e.rounds(cc, kk);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

double probe_det_exp(double x) { return det_exp(x); }
double probe_det_log(double x) { return det_log(x); }
int probe_count_eq(uint32_t w, uint32_t v) { return count_eq(w, v); }
uint32_t probe_sol_mask(uint32_t s) { return sol_mask(s); }
int probe_corner_edge(int b, int ax) { return corner_edge(b, ax); }
int probe_edge_start(int j) { return edge_start(j); }
int probe_edge_axis(int j) { return edge_axis(j); }
uint32_t probe_fdiv(uint32_t t, uint32_t d) { return fdiv(t, make_fastdiv(d)); }
uint32_t probe_uniform_int(uint32_t x, uint32_t n) { return uniform_int(x, n); }
double probe_u53(uint32_t a, uint32_t b) { return u53(a, b); }

int probe_shell_table(int L, float rc, float* d, int32_t* f, int cap) {
  std::vector<float> dd;
  std::vector<int32_t> ff;
  shell_table(L, rc, dd, ff);
  int n = (int)dd.size();
  for (int k = 0; k < n && k < cap; ++k) { d[k] = dd[k]; f[k] = ff[k]; }
  return n;
}

void probe_thresholds(float P, float T, float Jhb, float Jsig, float vhb, uint64_t* sig33,
                      uint64_t* eta9, uint64_t* sw2, double* jeff_beta) {
  const Thresholds t = make_thresholds(P, T, Jhb, Jsig, vhb);
  for (int i = 0; i < 33; ++i) sig33[i] = t.sigma[i];
  for (int i = 0; i < 9; ++i) eta9[i] = t.eta[i];
  sw2[0] = t.sw_hb; sw2[1] = t.sw_s;
  jeff_beta[0] = t.Jeff; jeff_beta[1] = t.beta;
}

double probe_vdw_fp64(double r, float rc, const float* d, const int32_t* f, int n) {
  return vdw_energy_fp64(r, rc, d, f, n);
}

// n host volume moves from step0 with the engine's code: out = r, V, E.
void probe_volume_chain(uint64_t seed, int L, float T, float P, float rc, int64_t step0, int64_t n,
                        double* out) {
  std::vector<float> d;
  std::vector<int32_t> f;
  shell_table(L, rc, d, f);
  const Thresholds t = make_thresholds(P, T, 0.5f, 0.08f, 0.6f);
  VolParams vp{seed, L, L * L * L, T, P, rc, (int32_t)d.size(), t.beta};
  VolState st;
  st.r_cell = 1.0001;
  st.V = (double)(L * L * L) * st.r_cell * st.r_cell * st.r_cell;
  st.E_vdw = vdw_energy_fp64(st.r_cell, rc, d.data(), f.data(), (int)d.size());
  st.accepted = 0;
  for (int64_t i = 0; i < n; ++i) {
    double u2;
    const double trial = volume_trial(st, vp, (uint64_t)(step0 + i), &u2);
    const double te = vdw_energy_fp64(trial, rc, d.data(), f.data(), (int)d.size());
    volume_accept(st, vp, trial, te, u2);
  }
  out[0] = st.r_cell; out[1] = st.V; out[2] = st.E_vdw; out[3] = st.accepted;
}
}
