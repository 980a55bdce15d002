/*
 * CPU oracle of the histogram reweighting program
 * (HistogramReweighting/histogram_reweighting_mixed_paramSI_norm_x.f90).
 *
 * TEST INFRASTRUCTURE ONLY: a plain C restatement of the Fortran program,
 * used by tests/ as the checker of the engine's reweighting library and of
 * the cvf_reweight executable.  The product never links or calls it.
 *
 * Layout follows the Fortran arrays: datos[(j * nb + i) * nb + w] is
 * datos_originales(j+1, i+1, w+1) (simulation j, energy bin i, density bin
 * w), log_P and memo_zero are [i * nb + w].
 */
#ifndef CVF_RW_ORACLE_H
#define CVF_RW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_rw_grid {
  int32_t L, N, n_sim, n_bin, n_bin2d, n_bin_mixed;
  double Emin, Dmin, dE, dD;  /* Emin, Densmin, tam_binE2D, tam_binDens2D */
  double *T, *P, *log_n;      /* K, bar, log(real(medidas)) */
  double *datos;              /* n_sim * nb * nb normalized histograms */
} ora_rw_grid;

/* Ranges, bin counts and normalized (E, rho) histograms of f90:116-283.
 * h_rho holds, simulation after simulation, counts[j] rows (h, rho) in
 * internal units.  Returns 0, or -1 on bad arguments. */
int ora_rw_grid_build(int32_t L, int32_t n_sim, const double *T_hat, const double *P_hat,
                      const int32_t *counts, const double *h_rho, double bin_factor, ora_rw_grid *g);
/* Same histograms from a caller-supplied datos grid (closed-form tests). */
int ora_rw_grid_from_datos(int32_t N, int32_t n_sim, int32_t nb, double Emin, double Dmin, double dE,
                           double dD, const double *T_K, const double *P_bar, const double *log_n,
                           const double *datos, ora_rw_grid *g);
void ora_rw_grid_free(ora_rw_grid *g);

/* Calcula_Histograma (f90:1003-1124): log_P of every cell at (temp K,
 * pres bar) given the constants, memo_zero = 1 on empty cells, log_Z.
 * Returns -1 when every cell is empty. */
int ora_rw_calcula(const ora_rw_grid *g, const double *ctes, double temp, double pres, double *log_P,
                   uint8_t *memo_zero, double *log_Z);

/* Constant iteration (f90:437-462): `steps` Jacobi sweeps from zero;
 * ctes_prev = the sweep before the last.  Returns 0, or 1 when some
 * |ctes - ctes_prev| > 1e-7 (not converged). */
int ora_rw_constants(const ora_rw_grid *g, int32_t steps, double *ctes, double *ctes_prev);

/* Normalized reweighted density on the grid (f90:933-955); 0 on empty cells. */
int ora_rw_distribution(const ora_rw_grid *g, const double *ctes, double temp, double pres, double *grid);

/* binder_energy_density (f90:2610-2662). */
void ora_rw_binder_ed(const ora_rw_grid *g, const double *grid, double *u_e, double *u_rho);

/* Calcula_Histograma_Rotado_normX (f90:1383-1619): order-parameter axis x
 * and density p (n_bin_mixed entries), with the first-stage mean `media`
 * and the second-stage scale `a1` (read by binder).  Returns 0, -1 when a
 * self-test (mean 0, norm 1, variance 1 within 1e-10) fails, -2 when a
 * cell falls outside the histogram. */
int ora_rw_normx(const ora_rw_grid *g, const double *grid, double alpha, double *x, double *p, double *media,
                 double *a1);

/* binder (f90:2580-2608). */
double ora_rw_binder(const double *x, const double *p, int32_t n, double media, double a1);

double ora_rw_ising2d(double z); /* f90:2218-2233 */
double ora_rw_ising3d(double z); /* f90:2235-2254 */
/* Ising_Peak (f90:1148-1224); returns -1 where the reference stops. */
int ora_rw_peaks(double *peak2d, double *peak3d);
/* ConfrontIsing, ZeroTest, HeightTest (f90:2042-2182):
 * out6 = {kl_2D, kl_3D, liu_2D, liu_3D, zero, height}. */
void ora_rw_scores(const double *x, const double *p, int32_t n, double peak2d, double peak3d, double out6[6]);

/* Err_Calcula_Histograma_Rotado (f90:1622-2039) with the distribution
 * `grid` at simulation k's (T, P) and that simulation's samples (E in
 * kJ/mol, rho in kg/m^3, `count` rows of two). */
int ora_rw_errors(const ora_rw_grid *g, const double *grid, double alpha, const double *samples, int32_t count,
                  double *error_hist, double *error_x);

/* The whole program over a directory holding `input` and the data_L*
 * files: writes Constants_L*, Binder_*, Confront_* and
 * Histogram_Reweighted_OP_* there and the summary lines to summary_path.
 * Returns 0 or the reference's exit status. */
int ora_rw_main(const char *dir, const char *summary_path);

#ifdef __cplusplus
}
#endif
#endif
