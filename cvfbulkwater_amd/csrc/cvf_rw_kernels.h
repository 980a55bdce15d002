// Device side of the histogram reweighting (include/cvf_reweight.h): the
// per-cell log-sum over simulations and the per-target reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvf {

constexpr int kRwBlock = 256;

// One simulation of the reweighting sums: beta_j = 1000/(0.2293 T_j) in
// mol/kJ, p_j = 100 P_j in kPa, lm_j = log(real(medidas_j)).
struct RwSim {
  double beta, pres, lm, pad;
};

// Occupied cells of the (e, rho) grid, structure of arrays in HBM.
struct RwCells {
  const double* en;  // e * N at the cell centre (kJ/mol)
  const double* vn;  // V * N = 0.018 N / rho (m^3/mol)
  const double* lh;  // log of the summed normalized histograms
  int64_t n;
};

struct RwTargets {
  const double2* bp;  // (beta, p) of each target
  int32_t n;
};

// Buffers of one reweighting pass over n targets.
struct RwWork {
  double* logp;  // [target][cell]: log_P, then (with dist) the density
  double* pmax;  // [target][block] partial maxima
  double* mx;    // [target] max log_P
  double* ps;    // [target][block][2] partial sums (truncated, full)
  double* lz;    // [target] log_Z
  double* norm;  // [target] sum of exp(log_P - log_Z) * de * drho
  int32_t nblk;
};

// log_P and log_Z of every target (Calcula_Histograma over all targets);
// with ctes_out, also writes log_Z(t) - log_Z(0) there (one Jacobi sweep
// of the constants); with dist, replaces log_P by the normalized density.
hipError_t launch_rw_pass(const RwCells& cells, const RwSim* sims, const double* ctes, int32_t n_sim,
                          const RwTargets& tg, const RwWork& w, double de, double drho, double* ctes_out,
                          bool dist, hipStream_t stream);

}  // namespace cvf
