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

// Grid geometry for the projections: cell = i << 16 | w (nb < 2^16), energy
// e_min + (i + 1/2) de (kJ/mol), density rho_min + (w + 1/2) drho (kg/m^3).
struct RwGeom {
  const uint32_t* cell;
  int64_t n;
  int32_t nb, nm;
  double e_min, de, rho_min, drho;
};

constexpr int kRwProjBlocks = 64;  // workgroups per target (or target x angle) of a projection

// Per target t (density dens[t][cell]): partial sums over each workgroup's
// cells of {v, v e^2, v e^4, v rho^2, v rho^4} -> part[t][block][5]
// (kRwProjBlocks blocks), for the Binder cumulants of the marginals.
hipError_t launch_rw_moments(const RwGeom& g, const double* dens, int32_t n_t, double* part, hipStream_t st);

// Per (target, angle) pair ta = t * n_a + a: the min and max of
// x = e sin(a) + rho cos(a) over the cells with density > 0 ->
// part[ta][block][2].  sc holds (sin, cos) per angle.
hipError_t launch_rw_xrange(const RwGeom& g, const double* dens, int32_t n_t, int32_t n_a, const double2* sc,
                            double* part, hipStream_t st);

// Per pair ta, the histogram of x over nm bins of width w from x0
// (binfo[ta] = (x0, w)), weighted by the density: part[ta][block][nm],
// summed in a fixed order (reproducible).  A cell with density > 0 outside
// the bins (and not within 1e-10 of x0, which goes to bin 0) sets *err.
hipError_t launch_rw_xhist(const RwGeom& g, const double* dens, int32_t n_t, int32_t n_a, const double2* sc,
                           const double2* binfo, double* part, uint32_t* err, hipStream_t st);

// out[ta][k] = sum over the kRwProjBlocks workgroups b, in order, of
// part[ta][b][k] (n_ta pairs, nm bins).
hipError_t launch_rw_xhist_sum(const double* part, int32_t n_ta, int32_t nm, double* out, hipStream_t st);

}  // namespace cvf
