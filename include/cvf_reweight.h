/*
 * Multiple-histogram reweighting of (energy, density) samples near the
 * liquid-liquid critical point: the engine's form of
 * HistogramReweighting/histogram_reweighting_mixed_paramSI_norm_x.f90.
 *
 * The Fortran program bins every simulation's (e, rho) samples on one 2D
 * grid, solves the reweighting constants by Jacobi iteration of
 * Calcula_Histograma (f90:437-462, 1003-1124), and then, for every (T, P)
 * of a scan, reweights the grid, projects it on the mixed order parameter
 * x = e sin(a) + rho cos(a) and compares it with the 3D Ising
 * distribution.  Here the occupied cells live in HBM as a list, the
 * per-cell log-sum over simulations runs as one HIP launch over all cells
 * and all targets (the constants' sweep or a batch of scan points), and a
 * scan's cumulants and order-parameter histograms are reduced on the device
 * (cvf_rw_scan); the per-target calls do the same on the host from a copy
 * of the densities.  The cvf_reweight executable drives these calls over
 * the reference's `input` file and data_L* files.
 *
 * Units follow the Fortran program: T_hat, P_hat internal; T in K, P in
 * bar; e in kJ/mol; rho in kg/m^3.  All functions return 0 (or a count)
 * or a negative cvf_status; cvf_last_error() holds the message.
 */
#ifndef CVF_REWEIGHT_H
#define CVF_REWEIGHT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cvf_reweight cvf_reweight;

/* Returned by cvf_rw_solve_constants when some constant moved by more than
 * 1e-7 in the last sweep (f90:465-471: "use a larger step"). */
#define CVF_RW_NOT_CONVERGED 1

typedef struct cvf_rw_info {
  int32_t n_sim;        /* simulations */
  int32_t N;            /* molecules, L^3 */
  int32_t n_bin;        /* 1D bins: int(sqrt(max count)/2) + 20 (f90:192-205) */
  int32_t n_bin2d;      /* 2D bins per axis: int(n_bin * 10 * bin_factor) (f90:207) */
  int32_t n_bin_mixed;  /* order-parameter bins: int(sqrt(n_bin2d)) (f90:926) */
  int32_t pad;
  int64_t n_occupied;   /* cells holding at least one sample */
  double e_min, rho_min; /* lower grid edges (Emin, Densmin) */
  double de, drho;       /* 2D bin widths (tam_binE2D, tam_binDens2D) */
} cvf_rw_info;

/* Histograms of n_sim simulations (f90:116-283).  h_rho holds, simulation
 * after simulation, counts[j] rows (h = H/N, rho = N/V) in internal units
 * (the awk '{print $3/N, N/$4}' of data.out); T_hat, P_hat the internal
 * temperatures and pressures; bin_factor the BIN_FCT of the driver script.
 * The reweighting kernels run on HIP device `device`. */
int cvf_rw_create(int32_t L, int32_t n_sim, const double *T_hat, const double *P_hat, const int64_t *counts,
                  const double *h_rho, double bin_factor, int32_t device, cvf_reweight **out);
void cvf_rw_destroy(cvf_reweight *rw);
int cvf_rw_get_info(const cvf_reweight *rw, cvf_rw_info *info);
/* Temperatures (K) and pressures (bar) of the simulations (f90:104-105). */
int cvf_rw_sim_points(const cvf_reweight *rw, double *T_K, double *P_bar);

/* Reweighting constants: `steps` Jacobi sweeps from zero (f90:437-462).
 * Writes the final constants and those of the sweep before; returns
 * CVF_RW_NOT_CONVERGED when some |difference| > 1e-7, else 0. */
int cvf_rw_solve_constants(cvf_reweight *rw, int32_t steps, double *ctes, double *ctes_prev);
/* Use constants read from a Constants_L<L> file (f90:386-434). */
int cvf_rw_set_constants(cvf_reweight *rw, const double *ctes);

/* Reweight to n_targets (T K, P bar) points in one launch sequence
 * (Calcula_Histograma + normalization, f90:933-955) and keep the
 * normalized densities for the calls below, target t = 0 .. n_targets-1.
 * log_z (may be NULL) receives log_Z of each target. */
int cvf_rw_reweight(cvf_reweight *rw, int32_t n_targets, const double *T_K, const double *P_bar, double *log_z);
/* Normalized density of target t on the full grid, n_bin2d x n_bin2d,
 * energy index major; 0 on empty cells. */
int cvf_rw_density(const cvf_reweight *rw, int32_t t, double *grid);
/* Binder cumulants 1 - <x^4>/(3<x^2>^2) of the energy and density
 * marginals (binder_energy_density, f90:2610-2662). */
int cvf_rw_binder_ed(const cvf_reweight *rw, int32_t t, double *u_e, double *u_rho);

/* Order-parameter distribution of target t at angle alpha
 * (Calcula_Histograma_Rotado_normX, f90:1383-1619): x and p (n_bin_mixed
 * entries) and stats3 = {media, a1, binder} where binder is the cumulant
 * of the binder subroutine (f90:2580-2608). */
int cvf_rw_order_parameter(const cvf_reweight *rw, int32_t t, double alpha, double *x, double *p, double *stats3);
/* Ising comparison of a distribution (x, p) (ConfrontIsing, ZeroTest,
 * HeightTest; f90:2042-2182): out6 = {kl_2D, kl_3D, liu_2D, liu_3D, zero,
 * height}. */
int cvf_rw_ising_scores(const double *x, const double *p, int32_t n, double *out6);
/* The scan of f90:928-1001 for n_targets (T K, P bar) points and n_angles
 * angles (n_targets x n_angles <= 4096) in one pass, projections on the
 * device: binder_ed[2t .. 2t+1] = the energy and density cumulants of
 * target t; for the pair ta = t * n_angles + a, x and p (n_bin_mixed each,
 * at ta * n_bin_mixed), stats3[3 ta ..] = {media, a1, binder} and
 * scores6[6 ta ..] as cvf_rw_order_parameter and cvf_rw_ising_scores give
 * them.  Sums run in an order fixed by the launch shape (reproducible run
 * to run).  The densities stay on the device: call cvf_rw_reweight before
 * the per-target functions. */
int cvf_rw_scan(cvf_reweight *rw, int32_t n_targets, const double *T_K, const double *P_bar, int32_t n_angles,
                const double *alpha, double *binder_ed, double *x, double *p, double *stats3, double *scores6);

/* Universal order-parameter distributions (f90:2218-2254). */
double cvf_rw_ising2d(double z);
double cvf_rw_ising3d(double z);

/* Error bars of the order-parameter histogram (Err_Calcula_Histograma_Rotado,
 * f90:1622-2039): target t must hold the density at simulation `sim`'s
 * (T, P); compares it with that simulation's own samples at angle alpha.
 * error_hist, error_x: n_bin_mixed entries. */
int cvf_rw_errors(const cvf_reweight *rw, int32_t t, int32_t sim, double alpha, double *error_hist,
                  double *error_x);

#ifdef __cplusplus
}
#endif
#endif
