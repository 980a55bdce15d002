/*
 * Thermodynamic post-processing of CVF state-point scans: the numerical
 * core of the reference's CVF_ThermodynamicAnalysis/ scripts, as a C ABI in
 * libcvf.so.  The executable built from cvfbulkwater_amd/csrc/
 * cvf_thermo_main.cpp drives it over a directory of runs the way
 * script_analisi.sh, filter_isobars.sh, calculadora.sh and conversor.sh do;
 * cvfbulkwater_amd.scan feeds it straight from a batched scan.
 *
 * All functions are host code and return 0 or a negative cvf_status.
 */
#ifndef CVF_THERMO_H
#define CVF_THERMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Value of K_T (and of its error) where it is not defined
 * (KTderivate.c:7). */
#define CVF_KT_NOT_VALUE (-10.0)

/* Means and population variances of the six observables of a data.out
 * series (columns e, h, v, N_HB, m, r_cell), accumulated in row order as
 * script_analisi.sh:35 does: out12 = {mean_e, var_e, mean_h, var_h, ...,
 * mean_r, var_r}, var = <x^2> - <x>^2.  `rows` holds n_rows rows of `stride`
 * doubles whose first six entries are e, h, v, N_HB, m, r_cell. */
int cvf_thermo_averages(const double *rows, int64_t n_rows, int32_t stride, double out12[12]);

/* Temperature derivative along an isobar, AlphaPderivate.c (mode 1:
 * alpha_P = (1/X) dX/dT) and CPderivate.c (mode 0: C_P = dX/dT), with their
 * error estimates: err(X_i) = sqrt(3 var_i / n_i), a third-derivative
 * truncation term and the propagated statistical term.  Points are sorted
 * by T first.  Rows 2 .. count-3 are written (count - 4 rows, 0 when
 * count < 5); returns the number of rows. */
int cvf_thermo_derivative(const double *T, const double *X, const double *var, const int32_t *n,
                          int32_t count, int32_t mode, double *T_out, double *d_out,
                          double *err_out);

/* Isothermal compressibility K_T = d ln(rho) / dP along isotherms
 * (KTderivate.c compresibilidad_log / error_compresibilidad_log): for each
 * point the neighbouring pressures on its isotherm (|dT| < 1e-8) give a
 * central (or one-sided) difference; points with rho < gas_rho, or a
 * neighbour below it, get CVF_KT_NOT_VALUE; the error adds a
 * third-derivative term.  The reference uses gas_rho = 500 (kg/m^3). */
int cvf_thermo_kt(const double *T, const double *P, const double *rho, const double *rho_err,
                  int32_t count, double gas_rho, double *KT, double *KT_err);

#ifdef __cplusplus
}
#endif
#endif
