// Host-side physics helpers: shell table and acceptance thresholds.
#include "cvf_host.h"

#include <math.h>
#include <stdlib.h>

namespace cvf {

void shell_table(int L, float r_cutoff, std::vector<float>& dist, std::vector<int32_t>& freq) {
  dist.clear();
  freq.clear();
  const int N = L * L * L, half = L / 2;
  for (int j = 1; j < N; ++j) {
    int dx = j % L, dy = (j / L) % L, dz = j / (L * L);
    if (dx > half) dx = L - dx;
    if (dy > half) dy = L - dy;
    if (dz > half) dz = L - dz;
    // distance_PBC (1323-1337): sqrt of the integer square sum, returned as float
    const float d = (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
    if (!(d < r_cutoff)) continue;
    size_t k = 0;
    for (; k < dist.size(); ++k)
      if (dist[k] == d) break;
    if (k == dist.size()) {
      dist.push_back(d);
      freq.push_back(1);
    } else {
      ++freq[k];
    }
  }
  for (auto& f : freq) f = (f / 2) * N;
}

static uint64_t p_to_threshold(double p) {
  if (!(p < 1.0)) return (uint64_t)1 << 32;
  if (p <= 0.0) return 0;
  return (uint64_t)floor(ldexp(p, 32));
}

Thresholds make_thresholds(float P, float T, float J_hb, float J_sig, float v_hb) {
  Thresholds t;
  const float jeff_f = J_hb - P * v_hb;
  t.Jeff = (double)jeff_f;
  t.beta = 1. / (double)T;
  for (int h = -1; h <= 1; ++h)
    for (int d = -5; d <= 5; ++d) {
      const double dE = -t.Jeff * h - (double)J_sig * d;
      t.sigma[(h + 1) * 11 + (d + 5)] = p_to_threshold(exp(-(t.beta * dE)));
    }
  for (int dn = -4; dn <= 4; ++dn) t.eta[dn + 4] = p_to_threshold(exp(t.beta * (t.Jeff * (double)dn)));
  t.sw_hb = p_to_threshold(1.0 - exp(-t.beta * fabs(t.Jeff)));
  t.sw_s = p_to_threshold(1.0 - exp(-t.beta * fabs((double)J_sig)));
  return t;
}

}  // namespace cvf
