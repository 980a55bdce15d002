// Multi-tau correlator (host): the Flag_correlation mode of the reference
// (gpu_water3D.cu:161-191, 2983-3128), after Ramirez, Sinha, Bhatt,
// Harmandaris and Schmid, J. Chem. Phys. 133, 154103 (2010).  Level k holds
// the last p values of the series block-averaged over m^k samples; level 0
// correlates lags 0..p-1, level k > 0 lags p/m..p-1 (in units of m^k).
// The update order (push to the next level before correlating) and the
// empty-slot sentinel follow the reference, since correlation.dat must be
// the same function of the input series.
#include <math.h>

#include <vector>

#include "cvf_api.h"

namespace {
constexpr int kLevels = 32, kP = 16, kM = 2;
constexpr double kEmpty = -2e10;
}  // namespace

struct cvf_correlator {
  double shift[kLevels][kP];
  double corr[kLevels][kP];
  long ncorr[kLevels][kP];
  double acc[kLevels];
  int nacc[kLevels];
  int ins[kLevels];
  int kmax = 0;
  double accval = 0;

  cvf_correlator() {
    for (int k = 0; k < kLevels; ++k) {
      for (int i = 0; i < kP; ++i) {
        shift[k][i] = kEmpty;
        corr[k][i] = 0;
        ncorr[k][i] = 0;
      }
      acc[k] = 0;
      nacc[k] = 0;
      ins[k] = 0;
    }
  }

  void add(double w, int k) {
    if (k == kLevels) return;
    if (k > kmax) kmax = k;
    shift[k][ins[k]] = w;
    if (k == 0) accval += w;
    acc[k] += w;
    if (++nacc[k] == kM) {
      add(acc[k] / kM, k + 1);
      acc[k] = 0;
      nacc[k] = 0;
    }
    const int i1 = ins[k];
    const int jmin = k == 0 ? 0 : kP / kM;
    for (int j = jmin; j < kP; ++j) {
      int i2 = i1 - j;
      if (i2 < 0) i2 += kP;
      if (shift[k][i2] > -1e10) {
        corr[k][j] += shift[k][i1] * shift[k][i2];
        ++ncorr[k][j];
      }
    }
    if (++ins[k] == kP) ins[k] = 0;
  }
};

extern "C" {

cvf_correlator* cvf_corr_create(void) { return new cvf_correlator(); }

void cvf_corr_add(cvf_correlator* c, double w) {
  if (c) c->add(w, 0);
}

int cvf_corr_result(cvf_correlator* c, double* t, double* f, int32_t cap) {
  if (!c || cap < 0 || (cap > 0 && (!t || !f))) return -1;
  if (c->ncorr[0][0] == 0) return 0;
  const double mean = c->accval / c->ncorr[0][0];
  const double aux = mean * mean;
  std::vector<double> tt, ff;
  for (int i = 0; i < kP; ++i)
    if (c->ncorr[0][i] > 0) {
      tt.push_back(i);
      ff.push_back(c->corr[0][i] / c->ncorr[0][i] - aux);
    }
  for (int k = 1; k < c->kmax; ++k)
    for (int i = kP / kM; i < kP; ++i)
      if (c->ncorr[k][i] > 0) {
        tt.push_back(i * pow((double)kM, k));
        ff.push_back(c->corr[k][i] / c->ncorr[k][i] - aux);
      }
  const int n = (int)tt.size() < cap ? (int)tt.size() : cap;
  for (int i = 0; i < n; ++i) {
    t[i] = tt[i];
    f[i] = ff[i] / ff[0];
  }
  return n;
}

void cvf_corr_destroy(cvf_correlator* c) { delete c; }

}  // extern "C"
