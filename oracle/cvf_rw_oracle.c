/*
 * CPU oracle of the histogram reweighting program: a plain C restatement of
 * HistogramReweighting/histogram_reweighting_mixed_paramSI_norm_x.f90.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the engine's reweighting
 * library and executable; never linked by the product).
 *
 * The Fortran program assigns default-real (single precision) literals to
 * real*8 variables and mixes them into real*8 expressions; SL() rounds
 * each such literal to float first, as the compiled program does.  Loops
 * run in the Fortran order (energy index outer, density index inner), so
 * every sum accumulates in the reference's order.
 *
 * Two places of the reference read memory they should not, and this
 * restatement takes the evident intent:
 *  - Err_Calcula_Histograma_Rotado converts the nearest simulation's
 *    samples with That(i), P(i) for a stale loop index i (f90:1902, 1936),
 *    out of the arrays' bounds; here with that simulation's own T and P.
 *  - ZeroTest leaves index_hist_min undefined when no bin has |x| < 1
 *    (f90:2045-2061); here that case takes the "1E3" branch.
 */
#include "cvf_rw_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SL(x) ((double)(float)(x))

/* unit conversions, f90:81-89 */
static double kTs(void) { return SL(140.5678); }
static double kTi(void) { return SL(185.4676); }
static double kPs(void) { return SL(4694.59); }
static double kPi(void) { return SL(-2111.5552); }
static double kRhos(void) { return SL(1527.3); }
static double kRhoi(void) { return SL(-23.1017); }
static double kHs(void) { return SL(11.272132); }
static double kHi(void) { return SL(9.028914); }

/* f90:161-162: rho in kg/m^3 and e = h - P v in kJ/mol */
static void convert(double h, double rho, double That, double P, double *E, double *D) {
  double d = kRhos() * rho + kRhoi();
  *D = d;
  *E = kHs() * h + kHi() * That - SL(1.8) * P / d;
}

static int fint(double q, int *out) { /* Fortran int(): truncation toward zero */
  if (!(fabs(q) < 1073741824.0)) return 0;
  *out = (int)q;
  return 1;
}

void ora_rw_grid_free(ora_rw_grid *g) {
  if (!g) return;
  free(g->T);
  free(g->P);
  free(g->log_n);
  free(g->datos);
  memset(g, 0, sizeof *g);
}

int ora_rw_grid_build(int32_t L, int32_t n_sim, const double *T_hat, const double *P_hat,
                      const int32_t *counts, const double *h_rho, double bin_factor, ora_rw_grid *g) {
  memset(g, 0, sizeof *g);
  if (L <= 0 || n_sim <= 0) return -1;
  g->L = L;
  g->N = L * L * L;
  g->n_sim = n_sim;
  g->T = malloc(sizeof(double) * n_sim);
  g->P = malloc(sizeof(double) * n_sim);
  g->log_n = malloc(sizeof(double) * n_sim);
  for (int i = 0; i < n_sim; ++i) {
    g->T[i] = T_hat[i] * kTs() + kTi(); /* f90:104-105 */
    g->P[i] = P_hat[i] * kPs() + kPi();
    g->log_n[i] = (double)logf((float)counts[i]); /* log(real(medidas)), f90:129 */
  }
  /* ranges, f90:139-190 */
  double Emin = 0, Emax = 0, Dmin = 0, Dmax = 0;
  int first = 1, mmax = 0;
  const double *row = h_rho;
  for (int i = 0; i < n_sim; ++i) {
    for (int j = 0; j < counts[i]; ++j, row += 2) {
      double E, D;
      convert(row[0], row[1], T_hat[i], g->P[i], &E, &D);
      if (first) {
        Emax = Emin = E;
        Dmax = Dmin = D;
        first = 0;
      } else {
        if (Emax < E) Emax = E;
        if (Emin > E) Emin = E;
        if (Dmax < D) Dmax = D;
        if (Dmin > D) Dmin = D;
      }
    }
    if (i == 0 || mmax < counts[i]) mmax = counts[i];
  }
  if (first) return -1;
  /* bins, f90:192-212 */
  double rangeE = Emax - Emin, rangeD = Dmax - Dmin;
  int n_bin = (int)(sqrtf((float)mmax) / 2.0f);
  if (n_bin < 1) return -1;
  double tE = rangeE / n_bin, tD = rangeD / n_bin;
  n_bin += 20;
  rangeE = rangeE + 20 * tE;
  Emin = Emin - 10 * tE;
  rangeD = rangeD + 20 * tD;
  Dmin = Dmin - 10 * tD;
  int nb = (int)((double)(n_bin * 10) * bin_factor);
  if (nb < 1) return -1;
  g->n_bin = n_bin;
  g->n_bin2d = nb;
  g->n_bin_mixed = (int)sqrtf((float)nb * 1.0f);
  g->Emin = Emin;
  g->Dmin = Dmin;
  g->dE = rangeE / nb;
  g->dD = rangeD / nb;
  /* histograms, f90:227-283 */
  g->datos = calloc((size_t)n_sim * nb * nb, sizeof(double));
  row = h_rho;
  for (int i = 0; i < n_sim; ++i) {
    double *d = g->datos + (size_t)i * nb * nb;
    double norma = counts[i];
    for (int j = 0; j < counts[i]; ++j, row += 2) {
      double E, D;
      int ie, id;
      convert(row[0], row[1], T_hat[i], g->P[i], &E, &D);
      int ok = fint((E - g->Emin) / g->dE, &ie) && fint((D - g->Dmin) / g->dD, &id);
      ie += 1;
      id += 1;
      if (ok && ie >= 1 && ie <= nb && id >= 1 && id <= nb)
        d[(size_t)(ie - 1) * nb + (id - 1)] += 1;
      else
        norma = norma - 1;
    }
    norma = norma * g->dE * g->dD;
    for (size_t c = 0; c < (size_t)nb * nb; ++c) d[c] = d[c] / norma;
  }
  return 0;
}

int ora_rw_grid_from_datos(int32_t N, int32_t n_sim, int32_t nb, double Emin, double Dmin, double dE, double dD,
                           const double *T_K, const double *P_bar, const double *log_n, const double *datos,
                           ora_rw_grid *g) {
  memset(g, 0, sizeof *g);
  if (N <= 0 || n_sim <= 0 || nb <= 0) return -1;
  g->N = N;
  g->n_sim = n_sim;
  g->n_bin = nb;
  g->n_bin2d = nb;
  g->n_bin_mixed = (int)sqrtf((float)nb);
  g->Emin = Emin;
  g->Dmin = Dmin;
  g->dE = dE;
  g->dD = dD;
  g->T = malloc(sizeof(double) * n_sim);
  g->P = malloc(sizeof(double) * n_sim);
  g->log_n = malloc(sizeof(double) * n_sim);
  memcpy(g->T, T_K, sizeof(double) * n_sim);
  memcpy(g->P, P_bar, sizeof(double) * n_sim);
  memcpy(g->log_n, log_n, sizeof(double) * n_sim);
  g->datos = malloc(sizeof(double) * (size_t)n_sim * nb * nb);
  memcpy(g->datos, datos, sizeof(double) * (size_t)n_sim * nb * nb);
  return 0;
}

int ora_rw_calcula(const ora_rw_grid *g, const double *ctes, double temp, double pres, double *log_P,
                   uint8_t *memo_zero, double *log_Z) {
  const int nb = g->n_bin2d, ns = g->n_sim;
  const double beta1 = 1000. / (SL(0.2293) * temp); /* f90:1013-1014 */
  const double pres1 = pres * 100.0;
  double *arg = malloc(sizeof(double) * ns);
  for (int i = 0; i < nb; ++i) {
    for (int w = 0; w < nb; ++w) {
      const size_t c = (size_t)i * nb + w;
      double hist = 0;
      for (int j = 0; j < ns; ++j) hist = hist + g->datos[(size_t)j * nb * nb + c];
      if (hist != 0) {
        double E = g->Emin + ((i + 1) - 0.5) * g->dE;
        double D = g->Dmin + ((w + 1) - 0.5) * g->dD;
        double V = SL(0.018) / D;
        E = E * g->N;
        V = V * g->N;
        const double log_hist = log(hist);
        int flag = 0;
        for (int j = 0; j < ns; ++j) {
          const double beta_in = 1000. / (SL(0.2293) * g->T[j]);
          const double pres_in = g->P[j] * 100.0;
          arg[j] = beta1 * (E + pres1 * V) - beta_in * (E + pres_in * V) - ctes[j] + g->log_n[j];
          if (fabs(arg[j]) > 10) flag = 1;
        }
        double xx = 0;
        if (flag == 0) {
          for (int j = 0; j < ns; ++j) xx = xx + exp(arg[j]);
          xx = log(xx);
        } else {
          double max_arg = arg[0];
          for (int j = 1; j < ns; ++j)
            if (max_arg < arg[j]) max_arg = arg[j];
          for (int j = 0; j < ns; ++j)
            if (max_arg - arg[j] < 10) xx = xx + exp(arg[j] - max_arg);
          if (xx != 0) xx = log(xx);
          xx = xx + max_arg;
        }
        log_P[c] = log_hist - xx;
        memo_zero[c] = 0;
      } else {
        memo_zero[c] = 1;
      }
    }
  }
  free(arg);
  /* f90:1084-1122 */
  long first = -1;
  for (size_t c = 0; c < (size_t)nb * nb; ++c)
    if (memo_zero[c] == 0) {
      first = (long)c;
      break;
    }
  if (first < 0) return -1;
  double maxlog = log_P[first];
  for (size_t c = 0; c < (size_t)nb * nb; ++c)
    if (memo_zero[c] == 0 && maxlog < log_P[c]) maxlog = log_P[c];
  double lz = 0;
  for (size_t c = 0; c < (size_t)nb * nb; ++c)
    if (memo_zero[c] == 0 && (maxlog - log_P[c]) < 10) lz = lz + exp(log_P[c] - maxlog);
  if (lz != 0) lz = log(lz);
  *log_Z = lz + maxlog;
  return 0;
}

int ora_rw_constants(const ora_rw_grid *g, int32_t steps, double *ctes, double *ctes_prev) {
  const int ns = g->n_sim;
  const size_t cells = (size_t)g->n_bin2d * g->n_bin2d;
  double *c = calloc(ns, sizeof(double)), *c2 = calloc(ns, sizeof(double));
  double *lp = malloc(sizeof(double) * cells);
  uint8_t *memo = malloc(cells);
  int rc = 0;
  for (int s = 0; s < steps && rc == 0; ++s) {
    memcpy(c, c2, sizeof(double) * ns);
    for (int k = 0; k < ns && rc == 0; ++k) {
      double lz;
      if (ora_rw_calcula(g, c, g->T[k], g->P[k], lp, memo, &lz) != 0) rc = -1;
      c2[k] = lz;
    }
    const double c0 = c2[0];
    for (int k = 0; k < ns; ++k) c2[k] = c2[k] - c0;
  }
  if (rc == 0)
    for (int k = 0; k < ns; ++k)
      if (fabs(c[k] - c2[k]) > SL(1E-07)) rc = 1;
  memcpy(ctes, c2, sizeof(double) * ns);
  if (ctes_prev) memcpy(ctes_prev, c, sizeof(double) * ns);
  free(c);
  free(c2);
  free(lp);
  free(memo);
  return rc;
}

int ora_rw_distribution(const ora_rw_grid *g, const double *ctes, double temp, double pres, double *grid) {
  const size_t cells = (size_t)g->n_bin2d * g->n_bin2d;
  double *lp = malloc(sizeof(double) * cells);
  uint8_t *memo = malloc(cells);
  double lz;
  int rc = ora_rw_calcula(g, ctes, temp, pres, lp, memo, &lz);
  if (rc == 0) {
    double norma = 0;
    for (size_t c = 0; c < cells; ++c) {
      grid[c] = 0;
      if (memo[c] == 0) {
        grid[c] = exp(lp[c] - lz);
        norma = norma + grid[c];
      }
    }
    norma = norma * g->dE * g->dD;
    for (size_t c = 0; c < cells; ++c)
      if (memo[c] == 0) grid[c] = grid[c] / norma;
  }
  free(lp);
  free(memo);
  return rc;
}

void ora_rw_binder_ed(const ora_rw_grid *g, const double *grid, double *u_e, double *u_rho) {
  const int nb = g->n_bin2d;
  double *he = calloc(nb, sizeof(double)), *hr = calloc(nb, sizeof(double));
  double norma = 0;
  for (int i = 0; i < nb; ++i)
    for (int w = 0; w < nb; ++w) {
      const double v = grid[(size_t)i * nb + w];
      he[i] = he[i] + v;
      hr[w] = hr[w] + v;
      norma = norma + v;
    }
  for (int i = 0; i < nb; ++i) {
    he[i] = he[i] / norma;
    hr[i] = hr[i] / norma;
  }
  double u2 = 0, u4 = 0;
  for (int i = 0; i < nb; ++i) {
    const double e = g->Emin + ((i + 1) - 0.5) * g->dE;
    const double m2 = e * e, m4 = m2 * m2;
    u2 = u2 + m2 * he[i];
    u4 = u4 + m4 * he[i];
  }
  *u_e = 1.0 - u4 / (3.0 * u2 * u2);
  u2 = 0;
  u4 = 0;
  for (int i = 0; i < nb; ++i) {
    const double r = g->Dmin + ((i + 1) - 0.5) * g->dD;
    const double m2 = r * r, m4 = m2 * m2;
    u2 = u2 + m2 * hr[i];
    u4 = u4 + m4 * hr[i];
  }
  *u_rho = 1.0 - u4 / (3.0 * u2 * u2);
  free(he);
  free(hr);
}

/* Rotate (f90:1126-1138): x = a sin(alpha) + b cos(alpha) */
static double rot_x(double E, double D, double alpha) { return E * sin(alpha) + D * cos(alpha); }

/* bin range of the rotated coordinate over cells with grid > 0, extended by
 * two bins on each side (f90:1396-1429) */
static void rotated_range(const ora_rw_grid *g, const double *grid, double alpha, int nbins, double *Xmin_out,
                          double *tX_out) {
  const int nb = g->n_bin2d;
  double Xmax = SL(-1E12), Xmin = SL(1E12);
  for (int i = 0; i < nb; ++i)
    for (int w = 0; w < nb; ++w)
      if (grid[(size_t)i * nb + w] > 0) {
        const double X = rot_x(g->Emin + ((i + 1) - 0.5) * g->dE, g->Dmin + ((w + 1) - 0.5) * g->dD, alpha);
        if (Xmax < X) Xmax = X;
        if (Xmin > X) Xmin = X;
      }
  double rangeX = Xmax - Xmin, tX = rangeX / nbins;
  rangeX = rangeX + 4 * tX;
  Xmin = Xmin - 2 * tX;
  tX = rangeX / nbins;
  *Xmin_out = Xmin;
  *tX_out = tX;
}

/* norm 1 / mean 0 / variance 1 with a uniform bin (f90:1484-1508) */
static void normalize_uniform(double *x, double *p, int n, double *tX, double *media_out) {
  double norma = 0;
  for (int i = 0; i < n; ++i) norma = norma + p[i];
  for (int i = 0; i < n; ++i) p[i] = p[i] / (norma * *tX);
  double media = 0;
  for (int i = 0; i < n; ++i) media = media + (x[i] * p[i] * *tX);
  for (int i = 0; i < n; ++i) x[i] = x[i] - media;
  norma = 0;
  double var = 0;
  for (int i = 0; i < n; ++i)
    if (p[i] > 0) {
      norma = norma + p[i] * *tX;
      var = var + (x[i] * x[i] * p[i] * *tX);
    }
  const double a1 = sqrt(norma / var), a2 = 1 / (norma * a1);
  for (int i = 0; i < n; ++i) x[i] = x[i] * a1;
  *tX = *tX * a1;
  for (int i = 0; i < n; ++i) p[i] = p[i] * a2;
  *media_out = media;
}

static int moments_ok(const double *x, const double *p, const double *w, double tX, int n) {
  double norma = 0, media = 0, var = 0;
  for (int i = 0; i < n; ++i)
    if (p[i] > 0) {
      const double b = w ? w[i] : tX;
      norma = norma + p[i] * b;
      media = media + (x[i] * p[i] * b);
      var = var + (x[i] * x[i] * p[i] * b);
    }
  var = var - media * media;
  return !(fabs(media) > SL(1E-10) || fabs(norma - 1) > SL(1E-10) || fabs(var - 1) > SL(1E-10));
}

int ora_rw_normx(const ora_rw_grid *g, const double *grid, double alpha, double *x, double *p, double *media_out,
                 double *a1_out) {
  const int nb = g->n_bin2d, nm = g->n_bin_mixed;
  double Xmin, tX;
  rotated_range(g, grid, alpha, nm, &Xmin, &tX);
  for (int i = 0; i < nm; ++i) p[i] = 0;
  for (int i = 0; i < nb; ++i)
    for (int w = 0; w < nb; ++w) {
      const double v = grid[(size_t)i * nb + w];
      if (v > 0) {
        const double X = rot_x(g->Emin + ((i + 1) - 0.5) * g->dE, g->Dmin + ((w + 1) - 0.5) * g->dD, alpha);
        int k;
        if (fint((X - Xmin) / tX, &k) && k + 1 >= 1 && k + 1 <= nm)
          p[k] = p[k] + v;
        else if (fabs(X - Xmin) < SL(1E-10))
          p[0] = p[0] + v;
        else
          return -2;
      }
    }
  for (int i = 0; i < nm; ++i) x[i] = Xmin + ((i + 1) - 0.5) * tX;
  double media;
  normalize_uniform(x, p, nm, &tX, &media);
  if (!moments_ok(x, p, NULL, tX, nm)) return -1;
  /* sine stretch of the axis and renormalization with per-bin widths,
   * f90:1530-1617 */
  const double b = SL(1.53), alh = SL(1.0), bt = SL(1.0), pi = SL(3.14159);
  for (int i = 0; i < nm; ++i) x[i] = alh * b * b * sin(pi * x[i] / b) / (pi * pi) + bt * x[i];
  double *rx = malloc(sizeof(double) * nm);
  rx[0] = x[1] - x[0];
  for (int i = 1; i < nm - 1; ++i) rx[i] = (x[i + 1] - x[i - 1]) / 2.0;
  rx[nm - 1] = x[nm - 1] - x[nm - 2];
  double norma = 0;
  for (int i = 0; i < nm; ++i) norma = norma + p[i] * rx[i];
  for (int i = 0; i < nm; ++i) p[i] = p[i] / norma;
  double media_rx = 0;
  for (int i = 0; i < nm; ++i) media_rx = media_rx + x[i] * p[i] * rx[i];
  for (int i = 0; i < nm; ++i) x[i] = x[i] - media_rx;
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
  const int ok = moments_ok(x, p, rx, 0, nm);
  free(rx);
  *media_out = media;
  *a1_out = a1;
  return ok ? 0 : -1;
}

double ora_rw_binder(const double *x, const double *p, int32_t n, double media, double a1) {
  double norma = 0;
  for (int i = 0; i < n; ++i) norma = norma + p[i];
  double u = 0, u2 = 0, u4 = 0;
  for (int i = 0; i < n; ++i) {
    const double op = (x[i] / a1) + media;
    const double m2 = op * op, m4 = m2 * m2;
    u = u + op * p[i] / norma;
    u2 = u2 + m2 * p[i] / norma;
    u4 = u4 + m4 * p[i] / norma;
  }
  (void)u;
  return 1.0 - u4 / (3.0 * u2 * u2);
}

double ora_rw_ising2d(double z) {
  const double a = SL(0.058137), b = SL(-2.68949), c = SL(-0.11235), d = SL(2.884162);
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4, z16 = z8 * z8;
  return exp(-(a * z16 + b * z * z + c * fabs(z) + d));
}

double ora_rw_ising3d(double z) {
  const double a = SL(0.196), c = SL(0.807), M0 = SL(1.141059404), A0 = SL(0.436591191);
  return A0 * exp(-(z * z / (M0 * M0) - 1) * (z * z / (M0 * M0) - 1) * (a * z * z / (M0 * M0) + c));
}

int ora_rw_peaks(double *peak2d, double *peak3d) {
  for (int which = 0; which < 2; ++which) {
    double (*f)(double) = which == 0 ? ora_rw_ising2d : ora_rw_ising3d;
    double xmin = SL(0.1), xmax = 2;
    double x1 = xmin + (xmax - xmin) / 3, x2 = xmin + 2 * (xmax - xmin) / 3;
    for (;;) {
      if (f(x1) > f(x2)) {
        xmax = x2;
      } else if (f(x2) > f(x1)) {
        xmin = x1;
      } else {
        return -1;
      }
      x1 = xmin + (xmax - xmin) / 3;
      x2 = xmin + 2 * (xmax - xmin) / 3;
      const int done = which == 0 ? fabs(x1 - x2) < SL(1E-06) : fabs(f(x1) - f(x2)) < SL(1E-06);
      if (done) {
        const double pk = f(x1) > f(x2) ? f(x1) : f(x2);
        if (which == 0)
          *peak2d = pk;
        else
          *peak3d = pk;
        break;
      }
    }
  }
  return 0;
}

void ora_rw_scores(const double *x, const double *p, int32_t n, double peak2d, double peak3d, double out6[6]) {
  double kl2 = 0, kl3 = 0, liu2 = 0, liu3 = 0;
  int n_liu = 0;
  for (int i = 0; i < n; ++i) {
    const double m = x[i], hist = p[i];
    if (hist != 0) {
      if (ora_rw_ising2d(m) > 0) kl2 = kl2 + (log(ora_rw_ising2d(m)) - log(hist)) * ora_rw_ising2d(m);
      if (ora_rw_ising3d(m) > 0) kl3 = kl3 + (log(ora_rw_ising3d(m)) - log(hist)) * ora_rw_ising3d(m);
      liu2 = liu2 + sqrt(hist) * fabs(hist - ora_rw_ising2d(m));
      liu3 = liu3 + sqrt(hist) * fabs(hist - ora_rw_ising3d(m));
      n_liu = n_liu + 1;
    }
  }
  liu2 = liu2 / ((peak2d - ora_rw_ising2d(0.0)) * n_liu);
  liu3 = liu3 / ((peak3d - ora_rw_ising3d(0.0)) * n_liu);
  const double k = (double)(6.0f / 0.8f);
  /* ZeroTest */
  double hist_min = SL(1E12), zero = SL(1E3);
  int imin = -1;
  for (int i = 0; i < n; ++i)
    if (x[i] > -1 && x[i] < 1 && p[i] < hist_min) {
      hist_min = p[i];
      imin = i;
    }
  if (imin >= 0) {
    const double dz = p[imin] - ora_rw_ising3d(0.0);
    zero = dz * dz;
    zero = zero * k * k;
    zero = zero + x[imin] * x[imin];
  }
  /* HeightTest */
  double h1 = -10, h2 = -10, height = SL(1E3);
  int i1 = -10, i2 = -10;
  for (int i = 0; i < n; ++i) {
    if (x[i] < 0) {
      if (p[i] > h1) {
        h1 = p[i];
        i1 = i;
      }
    } else if (p[i] > h2) {
      h2 = p[i];
      i2 = i;
    }
  }
  const double x1 = SL(-1.14), x2 = SL(1.14);
  if (i1 >= 0 && i2 >= 0) {
    const double d1 = p[i1] - ora_rw_ising3d(x1), d2 = p[i2] - ora_rw_ising3d(x2);
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
}

int ora_rw_errors(const ora_rw_grid *g, const double *grid, double alpha, const double *samples, int32_t count,
                  double *error_hist, double *error_x) {
  const int nb = g->n_bin2d, nm = g->n_bin_mixed, n1 = g->n_bin;
  /* (A) reweighted histogram along x, f90:1639-1747 */
  double Xmin, tX;
  rotated_range(g, grid, alpha, nm, &Xmin, &tX);
  double *hr = calloc(nm, sizeof(double)), *xr = malloc(sizeof(double) * nm);
  for (int i = 0; i < nb; ++i)
    for (int w = 0; w < nb; ++w) {
      const double X = rot_x(g->Emin + ((i + 1) - 0.5) * g->dE, g->Dmin + ((w + 1) - 0.5) * g->dD, alpha);
      int k;
      if (fint((X - Xmin) / tX, &k) && k + 1 >= 1 && k + 1 <= nm) hr[k] = hr[k] + grid[(size_t)i * nb + w];
    }
  for (int i = 0; i < nm; ++i) xr[i] = Xmin + ((i + 1) - 0.5) * tX;
  double media;
  normalize_uniform(xr, hr, nm, &tX, &media);
  int rc = moments_ok(xr, hr, NULL, tX, nm) ? 0 : -1;
  /* (B) the simulation's own samples, n_bin bins, f90:1881-2019 */
  double *ho = calloc(n1, sizeof(double)), *xo = malloc(sizeof(double) * n1);
  if (rc == 0) {
    double Xmax = SL(-1E12);
    Xmin = SL(1E12);
    for (int j = 0; j < count; ++j) {
      const double X = rot_x(samples[2 * j], samples[2 * j + 1], alpha);
      if (X > Xmax) Xmax = X;
      if (X < Xmin) Xmin = X;
    }
    double rangeX = Xmax - Xmin, to = rangeX / n1;
    rangeX = rangeX + 4 * to;
    Xmin = Xmin - 2 * to;
    to = rangeX / n1;
    double norma = count;
    for (int j = 0; j < count; ++j) {
      const double X = rot_x(samples[2 * j], samples[2 * j + 1], alpha);
      int k;
      if (fint((X - Xmin) / to, &k) && k + 1 >= 1 && k + 1 <= n1)
        ho[k] = ho[k] + 1;
      else
        norma = norma - 1;
    }
    for (int i = 0; i < n1; ++i) xo[i] = Xmin + ((i + 1) - 0.5) * to;
    for (int i = 0; i < n1; ++i) ho[i] = ho[i] / (norma * to);
    double m = 0;
    for (int i = 0; i < n1; ++i) m = m + (xo[i] * ho[i] * to);
    for (int i = 0; i < n1; ++i) xo[i] = xo[i] - m;
    norma = 0;
    double var = 0;
    for (int i = 0; i < n1; ++i)
      if (ho[i] > 0) {
        norma = norma + ho[i] * to;
        var = var + (xo[i] * xo[i] * ho[i] * to);
      }
    const double a1 = sqrt(norma / var), a2 = 1 / (norma * a1);
    for (int i = 0; i < n1; ++i) {
      xo[i] = xo[i] * a1;
      ho[i] = ho[i] * a2;
    }
    to = to * a1;
    if (!moments_ok(xo, ho, NULL, to, n1)) rc = -1;
  }
  if (rc == 0)
    for (int i = 0; i < nm; ++i) {
      double min_dist = SL(1E12);
      int min_idx = 0;
      for (int j = 0; j < n1; ++j) {
        const double dist = fabs(xo[j] - xr[i]);
        if (dist < min_dist) {
          min_dist = dist;
          min_idx = j;
        }
      }
      error_hist[i] = fabs(ho[min_idx] - hr[i]);
      error_x[i] = min_dist;
    }
  free(hr);
  free(xr);
  free(ho);
  free(xo);
  return rc;
}

/* ---------------- the program (f90:36-1001, 2256-2323) ---------------- */

static void p_label(char *buf, size_t n, double P, int strict) {
  /* pressure field of the output names: F7.1 / F6.1 / F5.1 */
  if (strict ? P < -1000.0 : P <= -1000.0)
    snprintf(buf, n, "%7.1f", P);
  else if (P < 0.0 || (strict ? P > 1000.0 : P >= 1000.0))
    snprintf(buf, n, "%6.1f", P);
  else
    snprintf(buf, n, "%5.1f", P);
}

static void wrow(FILE *f, const double *v, int n) {
  for (int i = 0; i < n; ++i) fprintf(f, " %.17g", v[i]);
  fprintf(f, "\n");
}

static void summary_line(FILE *f, const char *what, double val, double s, double T, double P, double B) {
  fprintf(f, " %s %.17g\n", what, val);
  fprintf(f, "%8s%9.6f%3s%9.5f%4s%6.1f\n%8s%9.6f\n", "S=", s, " T=", T, "K P=", P, "bar B=", B);
}

int ora_rw_main(const char *dir, const char *summary_path) {
  char path[4096], name[512];
  snprintf(path, sizeof path, "%s/input", dir);
  FILE *f = fopen(path, "r");
  if (!f) return 2;
  char key[64];
  int n_sim = 0, steps = 0, L = 0;
  double Tmin, Tmax, dT, Pmin, Pmax, dP, Smin, Smax, dS, bf;
  int ok = fscanf(f, "%63s %d", key, &n_sim) == 2 && fscanf(f, "%63s %d", key, &steps) == 2 &&
           fscanf(f, "%63s %d", key, &L) == 2 && fscanf(f, "%63s %lf", key, &Tmin) == 2 &&
           fscanf(f, "%63s %lf", key, &Tmax) == 2 && fscanf(f, "%63s %lf", key, &dT) == 2 &&
           fscanf(f, "%63s %lf", key, &Pmin) == 2 && fscanf(f, "%63s %lf", key, &Pmax) == 2 &&
           fscanf(f, "%63s %lf", key, &dP) == 2 && fscanf(f, "%63s %lf", key, &Smin) == 2 &&
           fscanf(f, "%63s %lf", key, &Smax) == 2 && fscanf(f, "%63s %lf", key, &dS) == 2 &&
           fscanf(f, "%63s %lf", key, &bf) == 2 && n_sim > 0;
  double *That = NULL, *Phat = NULL;
  int32_t *cnt = NULL;
  if (ok) {
    That = malloc(sizeof(double) * n_sim);
    Phat = malloc(sizeof(double) * n_sim);
    cnt = malloc(sizeof(int32_t) * n_sim);
    for (int i = 0; i < n_sim && ok; ++i) ok = fscanf(f, "%lf %lf %d", &That[i], &Phat[i], &cnt[i]) == 3;
  }
  fclose(f);
  if (!ok) return 2;
  if (Tmax <= Tmin || Pmax <= Pmin || Smax <= Smin) return 1;
  const double Tmin_SI = kTs() * Tmin + kTi(), Tmax_SI = kTs() * Tmax + kTi();
  /* samples */
  long total = 0;
  for (int i = 0; i < n_sim; ++i) total += cnt[i];
  double *hr = malloc(sizeof(double) * 2 * (total > 0 ? total : 1));
  long off = 0;
  for (int i = 0; i < n_sim && ok; ++i) {
    if (Phat[i] >= 0)
      snprintf(name, sizeof name, "data_L%d_T%9.7f_P%6.4f", L, That[i], Phat[i]);
    else
      snprintf(name, sizeof name, "data_L%d_T%9.7f_P%6.3f", L, That[i], Phat[i]);
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *d = fopen(path, "r");
    if (!d) {
      ok = 0;
      break;
    }
    for (int j = 0; j < cnt[i] && ok; ++j, off += 2) ok = fscanf(d, "%lf %lf", &hr[off], &hr[off + 1]) == 2;
    fclose(d);
  }
  ora_rw_grid g;
  if (!ok || ora_rw_grid_build(L, n_sim, That, Phat, cnt, hr, bf, &g) != 0) return 2;
  const int nb = g.n_bin2d, nm = g.n_bin_mixed;
  FILE *sum = fopen(summary_path, "w");
  /* constants, f90:386-495 */
  double *ctes = calloc(n_sim, sizeof(double)), *prev = calloc(n_sim, sizeof(double));
  snprintf(path, sizeof path, "%s/Constants_L%d", dir, L);
  FILE *cf = fopen(path, "r");
  int file_exists = cf != NULL, read_flag = 0;
  if (cf) {
    read_flag = 1;
    int i = 0;
    double tc, pc, vc;
    while (fscanf(cf, "%lf %lf %lf", &tc, &pc, &vc) == 3) {
      if (i >= n_sim) {
        read_flag = 0;
        break;
      }
      ctes[i] = vc;
      if (fabs(tc - That[i]) > SL(1E-10) || fabs(pc - Phat[i]) > SL(1E-10)) read_flag = 0;
      ++i;
    }
    if (i != n_sim) read_flag = 0;
    fclose(cf);
  }
  int rc = 0;
  if (!read_flag) {
    rc = ora_rw_constants(&g, steps, ctes, prev);
    if (rc != 0) rc = 1;
  }
  if (rc == 0) {
    fprintf(sum, " constants");
    for (int i = 0; i < n_sim; ++i) fprintf(sum, " %.17g", ctes[i]);
    fprintf(sum, "\n");
    if (!file_exists) {
      FILE *o = fopen(path, "w");
      for (int i = 0; i < n_sim; ++i) {
        const double v[3] = {That[i], Phat[i], ctes[i]};
        wrow(o, v, 3);
      }
      fclose(o);
    }
  }
  double peak2d = 0, peak3d = 0;
  if (rc == 0 && ora_rw_peaks(&peak2d, &peak3d) != 0) rc = 1;
  double *grid = malloc(sizeof(double) * (size_t)nb * nb);
  double *x = malloc(sizeof(double) * nm), *p = malloc(sizeof(double) * nm), *x_ = malloc(sizeof(double) * nm);
  FILE *f20 = NULL, *f21 = NULL, *f12 = NULL, *f13 = NULL;
  if (rc == 0) {
    char pl[32];
    p_label(pl, sizeof pl, g.P[0], 0);
    snprintf(path, sizeof path, "%s/Binder_L%d_P%sbar_Tm%9.5fK_TM%9.5fK_mA%8.6f_MA%8.6f.dat", dir, L, pl, Tmin_SI,
             Tmax_SI, Smin, Smax);
    f20 = fopen(path, "w");
    snprintf(path, sizeof path, "%s/Binder_L%d_P%sbar_Tm%9.5fK_TM%9.5fK_energy_density.dat", dir, L, pl, Tmin_SI,
             Tmax_SI);
    f21 = fopen(path, "w");
    snprintf(path, sizeof path, "%s/Confront_L%d_P%sbar_Tm%9.5fK_TM%9.5fK_mA%8.6f_MA%8.6fIsing3D_Liu.dat", dir, L,
             pl, Tmin_SI, Tmax_SI, Smin, Smax);
    f12 = fopen(path, "w");
    snprintf(path, sizeof path, "%s/Confront_L%d_P%sbar_Tm%9.5fK_TM%9.5fK_mA%8.6f_MA%8.6fIsing3D_KL.dat", dir, L,
             pl, Tmin_SI, Tmax_SI, Smin, Smax);
    f13 = fopen(path, "w");
  }
  /* minima: liu_2D, liu_3D, zero, height, zero+height: {value, S, T, P, B} */
  double best[5][5];
  for (int k = 0; k < 5; ++k) best[k][0] = SL(1E12);
  double alpha = 0, media = 0, a1 = 0;
  double P1 = Pmin;
  while (rc == 0 && P1 <= Pmax) {
    double T1 = Tmin;
    while (rc == 0 && T1 <= Tmax) {
      const double TK = T1 * kTs() + kTi(), PB = P1 * kPs() + kPi();
      if (ora_rw_distribution(&g, ctes, TK, PB, grid) != 0) {
        rc = 1;
        break;
      }
      double ue, ur;
      ora_rw_binder_ed(&g, grid, &ue, &ur);
      const double r21[4] = {TK, PB, ue, ur};
      wrow(f21, r21, 4);
      double S = Smin;
      while (S < Smax) {
        alpha = S;
        if (ora_rw_normx(&g, grid, alpha, x, p, &media, &a1) != 0) {
          rc = 1;
          break;
        }
        const double r20[4] = {TK, PB, S, ora_rw_binder(x, p, nm, media, a1)};
        wrow(f20, r20, 4);
        memcpy(x_, x, sizeof(double) * nm);
        double B = 1;
        while (B < SL(1.001)) {
          for (int i = 0; i < nm; ++i) x[i] = B * x_[i];
          double s6[6];
          ora_rw_scores(x, p, nm, peak2d, peak3d, s6);
          const double r12[4] = {TK, PB, S, s6[3]}, r13[4] = {TK, PB, S, s6[1]};
          wrow(f12, r12, 4);
          wrow(f13, r13, 4);
          const double vals[5] = {s6[2], s6[3], s6[4], s6[5], s6[4] + s6[5]};
          for (int k = 0; k < 5; ++k)
            if (best[k][0] > vals[k]) {
              best[k][0] = vals[k];
              best[k][1] = alpha;
              best[k][2] = TK;
              best[k][3] = PB;
              best[k][4] = B;
            }
          B = B + SL(0.01);
        }
        S = S + dS;
      }
      fprintf(f12, "\n");
      fprintf(f13, "\n");
      T1 = T1 + dT;
    }
    P1 = P1 + dP;
  }
  if (f20) fclose(f20);
  if (f21) fclose(f21);
  if (f12) fclose(f12);
  if (f13) fclose(f13);
  if (rc == 0) {
    static const char *what[5] = {"Liu 2D=", "Liu 3D=", "Zero Test=", "Height Test=", "Zero + Height Test="};
    for (int k = 0; k < 5; ++k) summary_line(sum, what[k], best[k][0], best[k][1], best[k][2], best[k][3], best[k][4]);
    /* errors at the simulation nearest to the liu_3D minimum, f90:931-965 */
    int err_k = 0;
    double dmin = 0;
    for (int k = 0; k < n_sim; ++k) {
      const double dd = (g.T[k] - best[1][2]) * (g.T[k] - best[1][2]) + (g.P[k] - best[1][3]) * (g.P[k] - best[1][3]);
      if (k == 0 || dd < dmin) {
        dmin = dd;
        err_k = k;
      }
    }
    double *eh = calloc(nm, sizeof(double)), *ex = calloc(nm, sizeof(double));
    long o0 = 0;
    for (int k = 0; k < err_k; ++k) o0 += 2L * cnt[k];
    double *smp = malloc(sizeof(double) * 2 * (cnt[err_k] > 0 ? cnt[err_k] : 1));
    for (int j = 0; j < cnt[err_k]; ++j)
      convert(hr[o0 + 2 * j], hr[o0 + 2 * j + 1], That[err_k], g.P[err_k], &smp[2 * j], &smp[2 * j + 1]);
    if (ora_rw_distribution(&g, ctes, g.T[err_k], g.P[err_k], grid) != 0 ||
        ora_rw_errors(&g, grid, best[1][1], smp, cnt[err_k], eh, ex) != 0)
      rc = 1;
    /* Print_Histogram at the liu_3D minimum, rotated by the last scanned
     * angle (the global alpha, f90:2285) */
    if (rc == 0 && ora_rw_distribution(&g, ctes, best[1][2], best[1][3], grid) == 0 &&
        ora_rw_normx(&g, grid, alpha, x, p, &media, &a1) == 0) {
      char pl[32];
      p_label(pl, sizeof pl, best[1][3], 1);
      snprintf(path, sizeof path, "%s/Histogram_Reweighted_OP_L%d_T%9.5f_P%s_A%8.6f_B%5.3f", dir, L, best[1][2], pl,
               best[1][1], best[1][4]);
      FILE *o = fopen(path, "w");
      fprintf(o, " #Data simulations (Temperature internal units): ");
      for (int i = 0; i < n_sim; ++i) fprintf(o, " %.17g", That[i]);
      fprintf(o, "\n");
      for (int i = 0; i < nm; ++i) {
        const double v[4] = {x[i] * best[1][4], p[i], ex[i], eh[i]};
        wrow(o, v, 4);
      }
      fclose(o);
      fprintf(sum, " normalization %.17g\n", a1);
    } else {
      rc = 1;
    }
    free(eh);
    free(ex);
    free(smp);
  }
  fclose(sum);
  free(grid);
  free(x);
  free(p);
  free(x_);
  free(ctes);
  free(prev);
  free(hr);
  free(That);
  free(Phat);
  free(cnt);
  ora_rw_grid_free(&g);
  return rc;
}
