// Host side of the histogram reweighting (include/cvf_reweight.h):
// sample conversion and binning, the occupied-cell list in HBM, the
// constants' Jacobi sweeps and batched reweighting on the device, the
// device-reduced scan, and the per-target projections, cumulants and Ising
// scores over the occupied cells on the host.
//
// Reference: HistogramReweighting/histogram_reweighting_mixed_paramSI_norm_x.f90.
// That program stores real*8 constants written as default-real literals
// (81-89, 1013, 1023, 1530-1533, 2224-2246) and mixes single-precision
// literals into FP64 expressions (162, 1048); sl() rounds each such
// literal to float first, as the compiled program does, so bin edges and
// weights agree with it to the last bit.
#include "cvf_reweight.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <utility>
#include <vector>

#include "cvf_api.h"
#include "cvf_host.h"
#include "cvf_rw_kernels.h"

using namespace cvf;

namespace {

constexpr double sl(float x) { return (double)x; }

// unit conversions (f90:81-89)
constexpr double kTs = sl(140.5678f), kTi = sl(185.4676f), kPs = sl(4694.59f), kPi = sl(-2111.5552f);
constexpr double kRhos = sl(1527.3f), kRhoi = sl(-23.1017f), kHs = sl(11.272132f), kHi = sl(9.028914f);

// rho in kg/m^3 and e = h - P v in kJ/mol (f90:161-162)
inline void to_si(double h, double rho, double T_hat, double P_bar, double& e, double& d) {
  d = kRhos * rho + kRhoi;
  e = kHs * h + kHi * T_hat - sl(1.8f) * P_bar / d;
}

// Fortran int(q) + 1 in [1, n]: k = int(q) in [0, n)
inline bool bin_of(double q, int n, int& k) {
  if (!(fabs(q) < 1073741824.0)) return false;
  k = (int)q;
  return k >= 0 && k < n;
}

// Rotate (f90:1126-1138): x = e sin(alpha) + rho cos(alpha)
inline double rot_x(double e, double rho, double alpha) { return e * sin(alpha) + rho * cos(alpha); }

#define HIPR(expr)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_error(CVF_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                                  \
  } while (0)

template <class T>
struct DMem {
  T* p = nullptr;
  DMem() = default;
  DMem(const DMem&) = delete;
  DMem& operator=(const DMem&) = delete;
  ~DMem() { free(); }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  hipError_t alloc(size_t n) {
    free();
    return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T));
  }
};

// Pinned host memory (hipHostMalloc), grown on demand: the densities'
// device-to-host copy then runs at the link's rate instead of through the
// runtime's staging of pageable memory.
template <class T>
struct HMem {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  HMem() = default;
  HMem(const HMem&) = delete;
  HMem& operator=(const HMem&) = delete;
  ~HMem() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t resize(size_t m) {
    if (m > cap) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = n = 0;
      const hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T), hipHostMallocDefault);
      if (e != hipSuccess) {
        p = nullptr;
        return e;
      }
      cap = m;
    }
    n = m;
    return hipSuccess;
  }
  T* data() { return p; }
  const T* data() const { return p; }
  size_t size() const { return n; }
};

// norm 1, mean 0, variance 1 with uniform bins of width *w (f90:1484-1508);
// returns the mean that was removed
double standardize(double* x, double* p, int n, double* w) {
  double norma = 0;
  for (int i = 0; i < n; ++i) norma = norma + p[i];
  for (int i = 0; i < n; ++i) p[i] = p[i] / (norma * *w);
  double media = 0;
  for (int i = 0; i < n; ++i) media = media + (x[i] * p[i] * *w);
  for (int i = 0; i < n; ++i) x[i] = x[i] - media;
  norma = 0;
  double var = 0;
  for (int i = 0; i < n; ++i)
    if (p[i] > 0) {
      norma = norma + p[i] * *w;
      var = var + (x[i] * x[i] * p[i] * *w);
    }
  const double a1 = sqrt(norma / var), a2 = 1 / (norma * a1);
  for (int i = 0; i < n; ++i) {
    x[i] = x[i] * a1;
    p[i] = p[i] * a2;
  }
  *w = *w * a1;
  return media;
}

// the reference's own check of the standardized histogram (f90:1510-1533):
// bins of width w[i] (or the uniform width w0)
bool standardized(const double* x, const double* p, const double* w, double w0, int n) {
  double norma = 0, media = 0, var = 0;
  for (int i = 0; i < n; ++i)
    if (p[i] > 0) {
      const double b = w ? w[i] : w0;
      norma = norma + p[i] * b;
      media = media + (x[i] * p[i] * b);
      var = var + (x[i] * x[i] * p[i] * b);
    }
  var = var - media * media;
  const double tol = sl(1E-10f);
  return !(fabs(media) > tol || fabs(norma - 1) > tol || fabs(var - 1) > tol);
}

// Ising_Peak (f90:1148-1224): ternary search of the maximum on [0.1, 2];
// the 2D search stops on |x1 - x2| < 1e-6, the 3D one on |f(x1) - f(x2)| < 1e-6
double ising_peak(double (*f)(double), bool by_x) {
  double lo = sl(0.1f), hi = 2;
  double x1 = lo + (hi - lo) / 3, x2 = lo + 2 * (hi - lo) / 3;
  for (;;) {
    if (f(x1) > f(x2))
      hi = x2;
    else if (f(x2) > f(x1))
      lo = x1;
    else
      return NAN;
    x1 = lo + (hi - lo) / 3;
    x2 = lo + 2 * (hi - lo) / 3;
    if (by_x ? fabs(x1 - x2) < sl(1E-06f) : fabs(f(x1) - f(x2)) < sl(1E-06f)) return f(x1) > f(x2) ? f(x1) : f(x2);
  }
}

// The order-parameter histogram p of nm bins of width w from xmin, as
// Calcula_Histograma_Rotado_normX has it after binning (f90:1466-1617):
// bin centres into x, norm / mean / variance, the sine stretch, the same
// with the stretched widths; stats3 = {media, a1, binder} (2580-2608).
int finish_order_parameter(double xmin, double w, int nm, double* x, double* p, double* stats3) {
  for (int i = 0; i < nm; ++i) x[i] = xmin + ((i + 1) - 0.5) * w;
  const double media = standardize(x, p, nm, &w);
  if (!standardized(x, p, nullptr, w, nm)) return set_error(CVF_ERANGE, "order-parameter moments off");
  // sine stretch of the axis (f90:1530-1534), then norm / mean / variance
  // again with the stretched bin widths (1538-1617)
  const double b = sl(1.53f), pi = sl(3.14159f);
  for (int i = 0; i < nm; ++i) x[i] = sl(1.0f) * b * b * sin(pi * x[i] / b) / (pi * pi) + sl(1.0f) * x[i];
  std::vector<double> rx(nm);
  rx[0] = x[1] - x[0];
  for (int i = 1; i < nm - 1; ++i) rx[i] = (x[i + 1] - x[i - 1]) / 2.0;
  rx[nm - 1] = x[nm - 1] - x[nm - 2];
  double norma = 0;
  for (int i = 0; i < nm; ++i) norma = norma + p[i] * rx[i];
  for (int i = 0; i < nm; ++i) p[i] = p[i] / norma;
  double mrx = 0;
  for (int i = 0; i < nm; ++i) mrx = mrx + x[i] * p[i] * rx[i];
  for (int i = 0; i < nm; ++i) x[i] = x[i] - mrx;
  norma = 0;
  double var = 0;
  for (int i = 0; i < nm; ++i)
    if (p[i] > 0) {
      norma = norma + p[i] * rx[i];
      var = var + (x[i] * x[i] * p[i] * rx[i]);
    }
  const double a1 = sqrt(norma / var), a2 = 1 / (norma * a1);
  for (int i = 0; i < nm; ++i) {
    x[i] = x[i] * a1;
    rx[i] = rx[i] * a1;
    p[i] = p[i] * a2;
  }
  if (!standardized(x, p, rx.data(), 0, nm)) return set_error(CVF_ERANGE, "stretched order-parameter moments off");
  if (stats3) {
    // binder (f90:2580-2608): cumulant of M = x / a1 + media
    double s = 0;
    for (int i = 0; i < nm; ++i) s = s + p[i];
    double u2 = 0, u4 = 0;
    for (int i = 0; i < nm; ++i) {
      const double m = (x[i] / a1) + media, m2 = m * m, m4 = m2 * m2;
      u2 = u2 + m2 * p[i] / s;
      u4 = u4 + m4 * p[i] / s;
    }
    stats3[0] = media;
    stats3[1] = a1;
    stats3[2] = 1.0 - u4 / (3.0 * u2 * u2);
  }
  return CVF_OK;
}

}  // namespace

struct cvf_reweight {
  int L = 0, N = 0, n_sim = 0, device = 0;
  std::vector<double> T_hat, T, P, log_n;
  std::vector<int64_t> counts, first;  // first[j]: index of simulation j's first sample
  std::vector<double> samples;         // (e, rho) of every sample, SI
  int n_bin = 0, nb = 0, nm = 0;
  double e_min = 0, rho_min = 0, de = 0, drho = 0;
  std::vector<int64_t> cell;  // occupied cells, i * nb + w ascending
  std::vector<double> ctes;
  bool have_ctes = false;

  hipStream_t stream = nullptr;
  DMem<double> en, vn, lh, c0, c1, logp, pmax, mx, ps, lz, norm;
  DMem<uint32_t> dcell;  // the occupied cells as i << 16 | w, for the projections
  DMem<RwSim> sims;
  DMem<double2> tgt;
  int32_t cap_t = 0, n_t = 0;
  HMem<double> dist;  // [target][cell] densities of the last reweighting (pinned)

  int32_t nblk() const { return (int32_t)((cell.size() + kRwBlock - 1) / kRwBlock); }
  RwCells cells() const { return RwCells{en.p, vn.p, lh.p, (int64_t)cell.size()}; }
  RwWork work() const { return RwWork{logp.p, pmax.p, mx.p, ps.p, lz.p, norm.p, nblk()}; }
  RwGeom geom() const { return RwGeom{dcell.p, (int64_t)cell.size(), nb, nm, e_min, de, rho_min, drho}; }

  int reserve_targets(int32_t n) {
    if (n <= cap_t) return CVF_OK;
    const size_t nc = cell.size(), b = (size_t)nblk();
    HIPR(logp.alloc((size_t)n * nc));
    HIPR(pmax.alloc((size_t)n * b));
    HIPR(ps.alloc((size_t)n * b * 2));
    HIPR(mx.alloc(n));
    HIPR(lz.alloc(n));
    HIPR(norm.alloc(n));
    HIPR(tgt.alloc(n));
    cap_t = n;
    return CVF_OK;
  }

  ~cvf_reweight() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" {

int cvf_rw_create(int32_t L, int32_t n_sim, const double* T_hat, const double* P_hat, const int64_t* counts,
                  const double* h_rho, double bin_factor, int32_t device, cvf_reweight** out) {
  if (!out || !T_hat || !P_hat || !counts || !h_rho) return set_error(CVF_EINVAL, "null argument");
  *out = nullptr;
  // n_sim <= 1024: the kernels hold 32 bytes per simulation in LDS
  if (L < 1 || L > 1024 || n_sim < 1 || n_sim > 1024) return set_error(CVF_EINVAL, "L=%d, n_sim=%d", L, n_sim);
  if (!(bin_factor > 0)) return set_error(CVF_EINVAL, "bin_factor must be > 0");
  int64_t total = 0, mmax = 0;
  for (int j = 0; j < n_sim; ++j) {
    if (counts[j] < 1 || counts[j] > 0x7FFFFFFF)
      return set_error(CVF_EINVAL, "simulation %d: %lld samples (need 1 .. 2^31-1)", j, (long long)counts[j]);
    total += counts[j];
    mmax = std::max<int64_t>(mmax, counts[j]);
  }
  auto rw = std::make_unique<cvf_reweight>();
  rw->L = L;
  rw->N = L * L * L;
  rw->n_sim = n_sim;
  rw->device = device;
  rw->T_hat.assign(T_hat, T_hat + n_sim);
  rw->counts.assign(counts, counts + n_sim);
  for (int j = 0; j < n_sim; ++j) {
    rw->T.push_back(T_hat[j] * kTs + kTi);  // f90:104-105
    rw->P.push_back(P_hat[j] * kPs + kPi);
    rw->log_n.push_back((double)logf((float)counts[j]));  // log(real(medidas)), f90:129
  }
  // SI samples and their ranges (f90:139-190)
  rw->samples.resize(2 * (size_t)total);
  double emin = 0, emax = 0, dmin = 0, dmax = 0;
  int64_t k = 0;
  for (int j = 0; j < n_sim; ++j) {
    rw->first.push_back(k);
    for (int64_t r = 0; r < counts[j]; ++r, ++k) {
      double e, d;
      to_si(h_rho[2 * k], h_rho[2 * k + 1], T_hat[j], rw->P[j], e, d);
      rw->samples[2 * k] = e;
      rw->samples[2 * k + 1] = d;
      if (k == 0) {
        emin = emax = e;
        dmin = dmax = d;
      } else {
        emax = std::max(emax, e);
        emin = std::min(emin, e);
        dmax = std::max(dmax, d);
        dmin = std::min(dmin, d);
      }
    }
  }
  // bins (f90:192-212)
  int n_bin = (int)(sqrtf((float)mmax) / 2.0f);
  if (n_bin < 1) return set_error(CVF_EINVAL, "need at least 4 samples in some simulation");
  double rangeE = emax - emin, rangeD = dmax - dmin;
  if (!(rangeE > 0) || !(rangeD > 0)) return set_error(CVF_EINVAL, "the samples span no energy or density range");
  const double tE = rangeE / n_bin, tD = rangeD / n_bin;
  n_bin += 20;
  rangeE = rangeE + 20 * tE;
  emin = emin - 10 * tE;
  rangeD = rangeD + 20 * tD;
  dmin = dmin - 10 * tD;
  const double nbd = (double)(n_bin * 10) * bin_factor;
  if (!(nbd >= 2) || nbd > 46340) return set_error(CVF_EINVAL, "n_bin2D = %g out of range [2, 46340]", nbd);
  const int nb = (int)nbd;
  rw->n_bin = n_bin;
  rw->nb = nb;
  rw->nm = (int)sqrtf((float)nb * 1.0f);
  rw->e_min = emin;
  rw->rho_min = dmin;
  rw->de = rangeE / nb;
  rw->drho = rangeD / nb;
  // per-simulation normalized histograms, summed over simulations in
  // simulation order (f90:227-283, 1020-1024), kept as a sorted cell list
  std::vector<std::pair<int64_t, double>> entries;  // (cell, count / norma_j)
  std::vector<int64_t> ids;
  k = 0;
  for (int j = 0; j < n_sim; ++j) {
    ids.clear();
    double norma = (double)counts[j];
    for (int64_t r = 0; r < counts[j]; ++r, ++k) {
      int ie, id;
      if (bin_of((rw->samples[2 * k] - emin) / rw->de, nb, ie) &&
          bin_of((rw->samples[2 * k + 1] - dmin) / rw->drho, nb, id))
        ids.push_back((int64_t)ie * nb + id);
      else
        norma = norma - 1;
    }
    norma = norma * rw->de * rw->drho;
    std::sort(ids.begin(), ids.end());
    for (size_t a = 0; a < ids.size();) {
      size_t b = a;
      while (b < ids.size() && ids[b] == ids[a]) ++b;
      entries.emplace_back(ids[a], (double)(b - a) / norma);
      a = b;
    }
  }
  std::stable_sort(entries.begin(), entries.end(),
                   [](const std::pair<int64_t, double>& a, const std::pair<int64_t, double>& b) {
                     return a.first < b.first;
                   });
  std::vector<double> hen, hvn, hlh;
  for (size_t a = 0; a < entries.size();) {
    double hist = 0;
    size_t b = a;
    for (; b < entries.size() && entries[b].first == entries[a].first; ++b) hist = hist + entries[b].second;
    const int64_t c = entries[a].first;
    const int i = (int)(c / nb), w = (int)(c % nb);
    const double e = emin + ((i + 1) - 0.5) * rw->de;
    const double d = dmin + ((w + 1) - 0.5) * rw->drho;
    rw->cell.push_back(c);
    hen.push_back(e * rw->N);  // f90:1026-1034
    hvn.push_back(sl(0.018f) / d * rw->N);
    hlh.push_back(log(hist));
    a = b;
  }
  if (rw->cell.empty()) return set_error(CVF_EINVAL, "no sample inside the histogram");
  // device state
  int ndev = 0;
  HIPR(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return set_error(CVF_EINVAL, "device %d of %d", device, ndev);
  HIPR(hipSetDevice(device));
  HIPR(hipStreamCreate(&rw->stream));
  const size_t nc = rw->cell.size();
  HIPR(rw->en.alloc(nc));
  HIPR(rw->vn.alloc(nc));
  HIPR(rw->lh.alloc(nc));
  HIPR(rw->c0.alloc(n_sim));
  HIPR(rw->c1.alloc(n_sim));
  HIPR(rw->sims.alloc(n_sim));
  HIPR(rw->dcell.alloc(nc));
  {
    // (energy row, density column) in 16 bits each: nb^2 < 2^31, so nb < 2^16
    std::vector<uint32_t> c32(nc);
    for (size_t k = 0; k < nc; ++k) {
      const uint64_t i = rw->cell[k] / (uint64_t)rw->nb, w = rw->cell[k] - i * (uint64_t)rw->nb;
      c32[k] = (uint32_t)(i << 16 | w);
    }
    HIPR(hipMemcpy(rw->dcell.p, c32.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  std::vector<RwSim> hs(n_sim);
  for (int j = 0; j < n_sim; ++j) hs[j] = RwSim{1000. / (sl(0.2293f) * rw->T[j]), rw->P[j] * 100.0, rw->log_n[j], 0};
  HIPR(hipMemcpy(rw->en.p, hen.data(), nc * sizeof(double), hipMemcpyHostToDevice));
  HIPR(hipMemcpy(rw->vn.p, hvn.data(), nc * sizeof(double), hipMemcpyHostToDevice));
  HIPR(hipMemcpy(rw->lh.p, hlh.data(), nc * sizeof(double), hipMemcpyHostToDevice));
  HIPR(hipMemcpy(rw->sims.p, hs.data(), n_sim * sizeof(RwSim), hipMemcpyHostToDevice));
  *out = rw.release();
  return CVF_OK;
}

void cvf_rw_destroy(cvf_reweight* rw) { delete rw; }

int cvf_rw_get_info(const cvf_reweight* rw, cvf_rw_info* info) {
  if (!rw || !info) return set_error(CVF_EINVAL, "null argument");
  *info = cvf_rw_info{rw->n_sim, rw->N, rw->n_bin, rw->nb, rw->nm, 0, (int64_t)rw->cell.size(),
                      rw->e_min, rw->rho_min, rw->de, rw->drho};
  return CVF_OK;
}

int cvf_rw_sim_points(const cvf_reweight* rw, double* T_K, double* P_bar) {
  if (!rw || !T_K || !P_bar) return set_error(CVF_EINVAL, "null argument");
  std::copy(rw->T.begin(), rw->T.end(), T_K);
  std::copy(rw->P.begin(), rw->P.end(), P_bar);
  return CVF_OK;
}

int cvf_rw_set_constants(cvf_reweight* rw, const double* ctes) {
  if (!rw || !ctes) return set_error(CVF_EINVAL, "null argument");
  rw->ctes.assign(ctes, ctes + rw->n_sim);
  rw->have_ctes = true;
  return CVF_OK;
}

int cvf_rw_solve_constants(cvf_reweight* rw, int32_t steps, double* ctes, double* ctes_prev) {
  if (!rw || !ctes) return set_error(CVF_EINVAL, "null argument");
  if (steps < 0) return set_error(CVF_EINVAL, "negative step count");
  HIPR(hipSetDevice(rw->device));
  if (int rc = rw->reserve_targets(rw->n_sim)) return rc;
  // the targets of a sweep are the simulations themselves
  std::vector<double2> bp(rw->n_sim);
  for (int j = 0; j < rw->n_sim; ++j) bp[j] = double2{1000. / (sl(0.2293f) * rw->T[j]), rw->P[j] * 100.0};
  HIPR(hipMemcpyAsync(rw->tgt.p, bp.data(), bp.size() * sizeof(double2), hipMemcpyHostToDevice, rw->stream));
  HIPR(hipMemsetAsync(rw->c0.p, 0, rw->n_sim * sizeof(double), rw->stream));
  HIPR(hipMemsetAsync(rw->c1.p, 0, rw->n_sim * sizeof(double), rw->stream));
  double *cur = rw->c0.p, *nxt = rw->c1.p;
  const RwTargets tg{rw->tgt.p, rw->n_sim};
  for (int s = 0; s < steps; ++s) {
    HIPR(launch_rw_pass(rw->cells(), rw->sims.p, cur, rw->n_sim, tg, rw->work(), rw->de, rw->drho, nxt, false,
                        rw->stream));
    std::swap(cur, nxt);
  }
  std::vector<double> prev(rw->n_sim);
  HIPR(hipMemcpyAsync(ctes, cur, rw->n_sim * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipMemcpyAsync(prev.data(), nxt, rw->n_sim * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipStreamSynchronize(rw->stream));
  if (ctes_prev) std::copy(prev.begin(), prev.end(), ctes_prev);
  rw->ctes.assign(ctes, ctes + rw->n_sim);
  rw->have_ctes = true;
  for (int j = 0; j < rw->n_sim; ++j)
    if (fabs(prev[j] - ctes[j]) > sl(1E-07f)) {
      set_error(CVF_RW_NOT_CONVERGED, "the constants have not converged: use more than %d steps", steps);
      return CVF_RW_NOT_CONVERGED;
    }
  return CVF_OK;
}

int cvf_rw_reweight(cvf_reweight* rw, int32_t n_targets, const double* T_K, const double* P_bar, double* log_z) {
  if (!rw || !T_K || !P_bar) return set_error(CVF_EINVAL, "null argument");
  if (n_targets < 1 || n_targets > 65535) return set_error(CVF_EINVAL, "n_targets=%d (1 .. 65535)", n_targets);
  if (!rw->have_ctes) return set_error(CVF_EINVAL, "the reweighting constants are not set");
  for (int t = 0; t < n_targets; ++t)
    if (!(T_K[t] > 0)) return set_error(CVF_EINVAL, "target %d: T = %g K", t, T_K[t]);
  HIPR(hipSetDevice(rw->device));
  if (int rc = rw->reserve_targets(n_targets)) return rc;
  std::vector<double2> bp(n_targets);
  for (int t = 0; t < n_targets; ++t) bp[t] = double2{1000. / (sl(0.2293f) * T_K[t]), P_bar[t] * 100.0};
  HIPR(hipMemcpyAsync(rw->tgt.p, bp.data(), bp.size() * sizeof(double2), hipMemcpyHostToDevice, rw->stream));
  HIPR(hipMemcpyAsync(rw->c0.p, rw->ctes.data(), rw->n_sim * sizeof(double), hipMemcpyHostToDevice, rw->stream));
  HIPR(launch_rw_pass(rw->cells(), rw->sims.p, rw->c0.p, rw->n_sim, RwTargets{rw->tgt.p, n_targets}, rw->work(),
                      rw->de, rw->drho, nullptr, true, rw->stream));
  const size_t nc = rw->cell.size();
  rw->n_t = 0;
  HIPR(rw->dist.resize((size_t)n_targets * nc));
  rw->n_t = n_targets;
  HIPR(hipMemcpyAsync(rw->dist.data(), rw->logp.p, rw->dist.size() * sizeof(double), hipMemcpyDeviceToHost,
                      rw->stream));
  std::vector<double> z(n_targets);
  HIPR(hipMemcpyAsync(z.data(), rw->lz.p, n_targets * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipStreamSynchronize(rw->stream));
  if (log_z) std::copy(z.begin(), z.end(), log_z);
  return CVF_OK;
}

int cvf_rw_density(const cvf_reweight* rw, int32_t t, double* grid) {
  if (!rw || !grid) return set_error(CVF_EINVAL, "null argument");
  if (t < 0 || t >= rw->n_t) return set_error(CVF_EINVAL, "target %d of %d", t, rw->n_t);
  std::fill(grid, grid + (size_t)rw->nb * rw->nb, 0.0);
  const double* v = rw->dist.data() + (size_t)t * rw->cell.size();
  for (size_t k = 0; k < rw->cell.size(); ++k) grid[rw->cell[k]] = v[k];
  return CVF_OK;
}

int cvf_rw_binder_ed(const cvf_reweight* rw, int32_t t, double* u_e, double* u_rho) {
  if (!rw || !u_e || !u_rho) return set_error(CVF_EINVAL, "null argument");
  if (t < 0 || t >= rw->n_t) return set_error(CVF_EINVAL, "target %d of %d", t, rw->n_t);
  const int nb = rw->nb;
  std::vector<double> he(nb, 0.0), hr(nb, 0.0);
  double norma = 0;
  const double* v = rw->dist.data() + (size_t)t * rw->cell.size();
  for (size_t k = 0; k < rw->cell.size(); ++k) {
    const int i = (int)(rw->cell[k] / nb), w = (int)(rw->cell[k] % nb);
    he[i] = he[i] + v[k];
    hr[w] = hr[w] + v[k];
    norma = norma + v[k];
  }
  // 1 - <x^4> / (3 <x^2>^2) over bin centres
  auto cumulant = [&](const std::vector<double>& h, double lo, double width) {
    double u2 = 0, u4 = 0;
    for (int i = 0; i < nb; ++i) {
      const double c = lo + ((i + 1) - 0.5) * width, m2 = c * c, m4 = m2 * m2, q = h[i] / norma;
      u2 = u2 + m2 * q;
      u4 = u4 + m4 * q;
    }
    return 1.0 - u4 / (3.0 * u2 * u2);
  };
  *u_e = cumulant(he, rw->e_min, rw->de);
  *u_rho = cumulant(hr, rw->rho_min, rw->drho);
  return CVF_OK;
}

int cvf_rw_order_parameter(const cvf_reweight* rw, int32_t t, double alpha, double* x, double* p, double* stats3) {
  if (!rw || !x || !p) return set_error(CVF_EINVAL, "null argument");
  if (t < 0 || t >= rw->n_t) return set_error(CVF_EINVAL, "target %d of %d", t, rw->n_t);
  const int nb = rw->nb, nm = rw->nm;
  if (nm < 3) return set_error(CVF_EINVAL, "n_bin_mixed = %d is too small", nm);
  const double* v = rw->dist.data() + (size_t)t * rw->cell.size();
  const size_t nc = rw->cell.size();
  std::vector<double> xc(nc);
  double xmax = sl(-1E12f), xmin = sl(1E12f);
  for (size_t k = 0; k < nc; ++k) {
    const int i = (int)(rw->cell[k] / nb), w = (int)(rw->cell[k] % nb);
    xc[k] = rot_x(rw->e_min + ((i + 1) - 0.5) * rw->de, rw->rho_min + ((w + 1) - 0.5) * rw->drho, alpha);
    if (v[k] > 0) {
      xmax = std::max(xmax, xc[k]);
      xmin = std::min(xmin, xc[k]);
    }
  }
  // two empty bins on each side (f90:1420-1429)
  double range = xmax - xmin, w = range / nm;
  range = range + 4 * w;
  xmin = xmin - 2 * w;
  w = range / nm;
  std::fill(p, p + nm, 0.0);
  for (size_t k = 0; k < nc; ++k) {
    if (!(v[k] > 0)) continue;
    int b;
    if (bin_of((xc[k] - xmin) / w, nm, b))
      p[b] = p[b] + v[k];
    else if (fabs(xc[k] - xmin) < sl(1E-10f))
      p[0] = p[0] + v[k];
    else
      return set_error(CVF_ERANGE, "order parameter %g outside the histogram", xc[k]);
  }
  return finish_order_parameter(xmin, w, nm, x, p, stats3);
}

int cvf_rw_scan(cvf_reweight* rw, int32_t n_targets, const double* T_K, const double* P_bar, int32_t n_angles,
                const double* alpha, double* binder_ed, double* x, double* p, double* stats3, double* scores6) {
  if (!rw || !T_K || !P_bar || !alpha || !binder_ed || !x || !p || !stats3 || !scores6)
    return set_error(CVF_EINVAL, "null argument");
  if (n_targets < 1 || n_angles < 1 || (int64_t)n_targets * n_angles > 4096)
    return set_error(CVF_EINVAL, "n_targets x n_angles = %d x %d (1 .. 4096 pairs)", n_targets, n_angles);
  if (!rw->have_ctes) return set_error(CVF_EINVAL, "the reweighting constants are not set");
  const int nm = rw->nm, nta = n_targets * n_angles;
  if (nm < 3) return set_error(CVF_EINVAL, "n_bin_mixed = %d is too small", nm);
  for (int t = 0; t < n_targets; ++t)
    if (!(T_K[t] > 0)) return set_error(CVF_EINVAL, "target %d: T = %g K", t, T_K[t]);
  HIPR(hipSetDevice(rw->device));
  if (int rc = rw->reserve_targets(n_targets)) return rc;
  // densities of every target, left on the device
  std::vector<double2> bp(n_targets);
  for (int t = 0; t < n_targets; ++t) bp[t] = double2{1000. / (sl(0.2293f) * T_K[t]), P_bar[t] * 100.0};
  HIPR(hipMemcpyAsync(rw->tgt.p, bp.data(), bp.size() * sizeof(double2), hipMemcpyHostToDevice, rw->stream));
  HIPR(hipMemcpyAsync(rw->c0.p, rw->ctes.data(), rw->n_sim * sizeof(double), hipMemcpyHostToDevice, rw->stream));
  HIPR(launch_rw_pass(rw->cells(), rw->sims.p, rw->c0.p, rw->n_sim, RwTargets{rw->tgt.p, n_targets}, rw->work(),
                      rw->de, rw->drho, nullptr, true, rw->stream));
  rw->n_t = 0;  // the host copy of the densities no longer matches the device
  const RwGeom g = rw->geom();
  std::vector<double2> sc(n_angles);
  for (int a = 0; a < n_angles; ++a) sc[a] = double2{sin(alpha[a]), cos(alpha[a])};  // as Rotate's sin / cos
  DMem<double2> dsc, dbin;
  DMem<double> pm, px, ph;
  DMem<uint32_t> derr;
  const int B = kRwProjBlocks;
  HIPR(dsc.alloc(n_angles));
  HIPR(dbin.alloc(nta));
  HIPR(pm.alloc((size_t)n_targets * B * 5));
  HIPR(px.alloc((size_t)nta * B * 2));
  HIPR(ph.alloc((size_t)nta * B * nm));
  HIPR(derr.alloc(1));
  HIPR(hipMemcpyAsync(dsc.p, sc.data(), sc.size() * sizeof(double2), hipMemcpyHostToDevice, rw->stream));
  HIPR(launch_rw_moments(g, rw->logp.p, n_targets, pm.p, rw->stream));
  HIPR(launch_rw_xrange(g, rw->logp.p, n_targets, n_angles, dsc.p, px.p, rw->stream));
  std::vector<double> hm((size_t)n_targets * B * 5), hx((size_t)nta * B * 2);
  HIPR(hipMemcpyAsync(hm.data(), pm.p, hm.size() * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipMemcpyAsync(hx.data(), px.p, hx.size() * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipStreamSynchronize(rw->stream));
  // Binder cumulants of the marginals: <x^2> = sum v x^2 / sum v
  for (int t = 0; t < n_targets; ++t) {
    double m[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < B; ++b)
      for (int k = 0; k < 5; ++k) m[k] += hm[((size_t)t * B + b) * 5 + k];
    binder_ed[2 * t] = 1.0 - (m[2] / m[0]) / (3.0 * (m[1] / m[0]) * (m[1] / m[0]));
    binder_ed[2 * t + 1] = 1.0 - (m[4] / m[0]) / (3.0 * (m[3] / m[0]) * (m[3] / m[0]));
  }
  // bins: the range over the occupied cells, two empty bins each side
  std::vector<double2> bins(nta);
  for (int ta = 0; ta < nta; ++ta) {
    double xmin = sl(1E12f), xmax = sl(-1E12f);
    for (int b = 0; b < B; ++b) {
      xmin = std::min(xmin, hx[2 * ((size_t)ta * B + b)]);
      xmax = std::max(xmax, hx[2 * ((size_t)ta * B + b) + 1]);
    }
    if (!(xmax >= xmin)) return set_error(CVF_ERANGE, "target %d: no cell of positive density", ta / n_angles);
    double range = xmax - xmin, w = range / nm;
    range = range + 4 * w;
    xmin = xmin - 2 * w;
    w = range / nm;
    bins[ta] = double2{xmin, w};
  }
  HIPR(hipMemcpyAsync(dbin.p, bins.data(), bins.size() * sizeof(double2), hipMemcpyHostToDevice, rw->stream));
  HIPR(hipMemsetAsync(derr.p, 0, sizeof(uint32_t), rw->stream));
  HIPR(launch_rw_xhist(g, rw->logp.p, n_targets, n_angles, dsc.p, dbin.p, ph.p, derr.p, rw->stream));
  // the workgroups' partial histograms summed on the device, in workgroup order
  DMem<double> phs;
  HIPR(phs.alloc((size_t)nta * nm));
  HIPR(launch_rw_xhist_sum(ph.p, nta, nm, phs.p, rw->stream));
  uint32_t err = 0;
  HIPR(hipMemcpyAsync(p, phs.p, (size_t)nta * nm * sizeof(double), hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipMemcpyAsync(&err, derr.p, sizeof err, hipMemcpyDeviceToHost, rw->stream));
  HIPR(hipStreamSynchronize(rw->stream));
  if (err) return set_error(CVF_ERANGE, "an order parameter fell outside its histogram");
  for (int ta = 0; ta < nta; ++ta) {
    double* xo = x + (size_t)ta * nm;
    double* po = p + (size_t)ta * nm;
    if (int rc = finish_order_parameter(bins[ta].x, bins[ta].y, nm, xo, po, stats3 + 3 * ta)) return rc;
    if (int rc = cvf_rw_ising_scores(xo, po, nm, scores6 + 6 * ta)) return rc;
  }
  return CVF_OK;
}

double cvf_rw_ising2d(double z) {
  // Liu-Panagiotopoulos fit (f90:2218-2233)
  const double a = sl(0.058137f), b = sl(-2.68949f), c = sl(-0.11235f), d = sl(2.884162f);
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4, z16 = z8 * z8;
  return exp(-(a * z16 + b * z * z + c * fabs(z) + d));
}

double cvf_rw_ising3d(double z) {
  // Tsypin form with unit norm and variance (f90:2235-2254)
  const double a = sl(0.196f), c = sl(0.807f), M0 = sl(1.141059404f), A0 = sl(0.436591191f);
  const double q = z * z / (M0 * M0) - 1;
  return A0 * exp(-q * q * (a * z * z / (M0 * M0) + c));
}

int cvf_rw_ising_scores(const double* x, const double* p, int32_t n, double* out6) {
  if (!x || !p || !out6 || n < 1) return set_error(CVF_EINVAL, "bad arguments");
  static const double peak2d = ising_peak(cvf_rw_ising2d, true), peak3d = ising_peak(cvf_rw_ising3d, false);
  double kl2 = 0, kl3 = 0, liu2 = 0, liu3 = 0;
  int n_liu = 0;
  for (int i = 0; i < n; ++i) {
    if (p[i] == 0) continue;
    const double f2 = cvf_rw_ising2d(x[i]), f3 = cvf_rw_ising3d(x[i]);
    if (f2 > 0) kl2 = kl2 + (log(f2) - log(p[i])) * f2;
    if (f3 > 0) kl3 = kl3 + (log(f3) - log(p[i])) * f3;
    liu2 = liu2 + sqrt(p[i]) * fabs(p[i] - f2);
    liu3 = liu3 + sqrt(p[i]) * fabs(p[i] - f3);
    ++n_liu;
  }
  liu2 = liu2 / ((peak2d - cvf_rw_ising2d(0.0)) * n_liu);
  liu3 = liu3 / ((peak3d - cvf_rw_ising3d(0.0)) * n_liu);
  const double k = (double)(6.0f / 0.8f);  // plot aspect of the reference's figures
  // ZeroTest: the lowest bin with |x| < 1 against Ising3D(0)
  int imin = -1;
  double hmin = sl(1E12f), zero = sl(1E3f);
  for (int i = 0; i < n; ++i)
    if (x[i] > -1 && x[i] < 1 && p[i] < hmin) {
      hmin = p[i];
      imin = i;
    }
  if (imin >= 0) {
    const double d = p[imin] - cvf_rw_ising3d(0.0);
    zero = d * d;
    zero = zero * k * k;
    zero = zero + x[imin] * x[imin];
  }
  // HeightTest: the highest bin on each side against the maxima at -/+1.14
  int i1 = -1, i2 = -1;
  double h1 = -10, h2 = -10, height = sl(1E3f);
  for (int i = 0; i < n; ++i) {
    if (x[i] < 0) {
      if (p[i] > h1) h1 = p[i], i1 = i;
    } else if (p[i] > h2) {
      h2 = p[i], i2 = i;
    }
  }
  if (i1 >= 0 && i2 >= 0) {
    const double x1 = sl(-1.14f), x2 = sl(1.14f);
    const double d1 = p[i1] - cvf_rw_ising3d(x1), d2 = p[i2] - cvf_rw_ising3d(x2);
    height = d1 * d1 * k * k;
    height = height + d2 * d2 * k * k;
    height = height + (x[i1] - x1) * (x[i1] - x1);
    height = height + (x[i2] - x2) * (x[i2] - x2);
  }
  out6[0] = kl2;
  out6[1] = kl3;
  out6[2] = liu2;
  out6[3] = liu3;
  out6[4] = zero;
  out6[5] = height;
  return CVF_OK;
}

int cvf_rw_errors(const cvf_reweight* rw, int32_t t, int32_t sim, double alpha, double* error_hist,
                  double* error_x) {
  if (!rw || !error_hist || !error_x) return set_error(CVF_EINVAL, "null argument");
  if (t < 0 || t >= rw->n_t) return set_error(CVF_EINVAL, "target %d of %d", t, rw->n_t);
  if (sim < 0 || sim >= rw->n_sim) return set_error(CVF_EINVAL, "simulation %d of %d", sim, rw->n_sim);
  const int nb = rw->nb, nm = rw->nm, n1 = rw->n_bin;
  const size_t nc = rw->cell.size();
  const double* v = rw->dist.data() + (size_t)t * nc;
  // (A) the reweighted density along x (f90:1639-1747)
  std::vector<double> xc(nc), hr(nm, 0.0), xr(nm);
  double xmax = sl(-1E12f), xmin = sl(1E12f);
  for (size_t k = 0; k < nc; ++k) {
    const int i = (int)(rw->cell[k] / nb), w = (int)(rw->cell[k] % nb);
    xc[k] = rot_x(rw->e_min + ((i + 1) - 0.5) * rw->de, rw->rho_min + ((w + 1) - 0.5) * rw->drho, alpha);
    if (v[k] > 0) {
      xmax = std::max(xmax, xc[k]);
      xmin = std::min(xmin, xc[k]);
    }
  }
  double range = xmax - xmin, w = range / nm;
  range = range + 4 * w;
  xmin = xmin - 2 * w;
  w = range / nm;
  for (size_t k = 0; k < nc; ++k) {
    int b;
    if (bin_of((xc[k] - xmin) / w, nm, b)) hr[b] = hr[b] + v[k];
  }
  for (int i = 0; i < nm; ++i) xr[i] = xmin + ((i + 1) - 0.5) * w;
  standardize(xr.data(), hr.data(), nm, &w);
  if (!standardized(xr.data(), hr.data(), nullptr, w, nm))
    return set_error(CVF_ERANGE, "reweighted order-parameter moments off");
  // (B) the simulation's own samples in n_bin bins (f90:1881-2019)
  const double* s = rw->samples.data() + 2 * rw->first[sim];
  const int64_t cnt = rw->counts[sim];
  xmax = sl(-1E12f);
  xmin = sl(1E12f);
  for (int64_t j = 0; j < cnt; ++j) {
    const double X = rot_x(s[2 * j], s[2 * j + 1], alpha);
    xmax = std::max(xmax, X);
    xmin = std::min(xmin, X);
  }
  range = xmax - xmin;
  double wo = range / n1;
  range = range + 4 * wo;
  xmin = xmin - 2 * wo;
  wo = range / n1;
  std::vector<double> ho(n1, 0.0), xo(n1);
  for (int64_t j = 0; j < cnt; ++j) {
    int b;
    if (bin_of((rot_x(s[2 * j], s[2 * j + 1], alpha) - xmin) / wo, n1, b)) ho[b] = ho[b] + 1;
  }
  for (int i = 0; i < n1; ++i) xo[i] = xmin + ((i + 1) - 0.5) * wo;
  // the counts sum exactly to the in-range count, the reference's norm
  standardize(xo.data(), ho.data(), n1, &wo);
  if (!standardized(xo.data(), ho.data(), nullptr, wo, n1))
    return set_error(CVF_ERANGE, "sampled order-parameter moments off");
  // nearest sampled bin of every reweighted bin
  for (int i = 0; i < nm; ++i) {
    double dmin = sl(1E12f);
    int jmin = 0;
    for (int j = 0; j < n1; ++j) {
      const double d = fabs(xo[j] - xr[i]);
      if (d < dmin) dmin = d, jmin = j;
    }
    error_hist[i] = fabs(ho[jmin] - hr[i]);
    error_x[i] = dmin;
  }
  return CVF_OK;
}

}  // extern "C"
