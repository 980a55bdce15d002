// Thermodynamic post-processing: the numerical core of the reference's
// CVF_ThermodynamicAnalysis/ (script_analisi.sh, AlphaPderivate.c,
// CPderivate.c, KTderivate.c), host code behind include/cvf_thermo.h.
#include "cvf_thermo.h"

#include <math.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "cvf_api.h"

extern "C" {

int cvf_thermo_averages(const double* rows, int64_t n_rows, int32_t stride, double out12[12]) {
  if (!rows || !out12 || n_rows <= 0 || stride < 6) return CVF_EINVAL;
  // running sums in row order (script_analisi.sh:35)
  double s[6] = {0, 0, 0, 0, 0, 0}, s2[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t r = 0; r < n_rows; ++r) {
    const double* x = rows + r * stride;
    for (int k = 0; k < 6; ++k) {
      s[k] += x[k];
      s2[k] += x[k] * x[k];
    }
  }
  const double nr = (double)n_rows;
  for (int k = 0; k < 6; ++k) {
    out12[2 * k] = s[k] / nr;
    out12[2 * k + 1] = s2[k] / nr - (s[k] / nr) * (s[k] / nr);
  }
  return CVF_OK;
}

int cvf_thermo_derivative(const double* T, const double* X, const double* var, const int32_t* n,
                          int32_t count, int32_t mode, double* T_out, double* d_out, double* err_out) {
  if (!T || !X || !var || !n || count < 0 || (mode != 0 && mode != 1)) return CVF_EINVAL;
  if (count < 5) return 0;
  if (!T_out || !d_out || !err_out) return CVF_EINVAL;
  // ascending T (the reference's selection sort keeps the first of equal T)
  std::vector<int> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return T[a] < T[b]; });
  std::vector<double> t(count), x(count), ex(count), d(count);
  for (int j = 0; j < count; ++j) {
    t[j] = T[ord[j]];
    x[j] = X[ord[j]];
    // err = sqrt(3 sigma^2 / n): sampling interval ~ correlation time
    ex[j] = sqrt((var[ord[j]] / n[ord[j]]) * 3);
  }
  for (int j = 0; j < count; ++j) {
    if (j == 0)
      d[j] = (x[j + 1] - x[j]) / (t[j + 1] - t[j]);
    else if (j == count - 1)
      d[j] = (x[j] - x[j - 1]) / (t[j] - t[j - 1]);
    else
      d[j] = (x[j + 1] - x[j - 1]) / (t[j + 1] - t[j - 1]);
    if (mode == 1) d[j] *= 1. / x[j];
  }
  int m = 0;
  for (int j = 2; j < count - 2; ++j) {
    const double dT = (t[j + 1] - t[j - 1]) / 2.0;
    // truncation error of the central difference from the third derivative
    double errDer = fabs(x[j + 2] / 2.0 - x[j + 1] + x[j - 1] - x[j - 2] / 2.0) / (dT * dT * dT);
    errDer *= dT * dT / 6.0;
    const double errEst = (ex[j + 1] + ex[j - 1]) / (2.0 * dT);
    double err;
    if (mode == 1)
      err = sqrt(d[j] * d[j] * ex[j] * ex[j] / (x[j] * x[j]) + errDer * errDer / (x[j] * x[j]) +
                 errEst * errEst / (x[j] * x[j]));
    else
      err = sqrt(errDer * errDer + errEst * errEst);
    T_out[m] = t[j];
    d_out[m] = d[j];
    err_out[m] = err;
    ++m;
  }
  return m;
}

namespace {

// nearest pressures above and below point j on its isotherm (and, for the
// error, the next ones beyond those)
struct Neighbours {
  int up = -1, dn = -1, up2 = -1, dn2 = -1;
  double Pup = 1e12, Pdn = -1e12, Pup2 = 1e12, Pdn2 = -1e12;
};

Neighbours neighbours(const double* T, const double* P, int count, int j) {
  Neighbours nb;
  for (int i = 0; i < count; ++i) {
    if (i == j || !(fabs(T[i] - T[j]) < 1e-8)) continue;
    if (P[i] > P[j] && P[i] < nb.Pup) {
      nb.Pup = P[i];
      nb.up = i;
    }
    if (P[i] < P[j] && P[i] > nb.Pdn) {
      nb.Pdn = P[i];
      nb.dn = i;
    }
  }
  for (int i = 0; i < count; ++i) {
    if (i == j || !(fabs(T[i] - T[j]) < 1e-8)) continue;
    if (P[i] > nb.Pup && P[i] < nb.Pup2) {
      nb.Pup2 = P[i];
      nb.up2 = i;
    }
    if (P[i] < nb.Pdn && P[i] > nb.Pdn2) {
      nb.Pdn2 = P[i];
      nb.dn2 = i;
    }
  }
  return nb;
}

}  // namespace

int cvf_thermo_kt(const double* T, const double* P, const double* rho, const double* rho_err,
                  int32_t count, double gas_rho, double* KT, double* KT_err) {
  if (!T || !P || !rho || !rho_err || !KT || !KT_err || count < 0) return CVF_EINVAL;
  const double NV = CVF_KT_NOT_VALUE;
  std::vector<Neighbours> nbs(count);
  // K_T = d ln(rho)/dP (compresibilidad_log, KTderivate.c:429-481)
  for (int j = 0; j < count; ++j) {
    const Neighbours nb = nbs[j] = neighbours(T, P, count, j);
    const double dens = rho[j], pres = P[j];
    if (dens < gas_rho || (nb.up == -1 && nb.dn == -1)) {
      KT[j] = KT_err[j] = NV;
    } else if (nb.up == -1) {  // highest pressure of the isotherm
      if (rho[nb.dn] < gas_rho) {
        KT[j] = KT_err[j] = NV;
      } else {
        KT[j] = (log(dens) - log(rho[nb.dn])) / (pres - nb.Pdn);
        KT_err[j] = fabs(rho_err[j] / dens + rho_err[nb.dn] / rho[nb.dn]) / (pres - nb.Pdn);
      }
    } else if (nb.dn == -1) {  // lowest pressure of the isotherm
      if (rho[nb.up] < gas_rho) {
        KT[j] = KT_err[j] = NV;
      } else {
        KT[j] = (log(rho[nb.up]) - log(dens)) / (nb.Pup - pres);
        KT_err[j] = fabs(rho_err[j] / dens + rho_err[nb.up] / rho[nb.up]) / (nb.Pup - pres);
      }
    } else {
      if (rho[nb.up] < gas_rho || rho[nb.dn] < gas_rho) {
        KT[j] = KT_err[j] = NV;
      } else {
        KT[j] = (log(rho[nb.up]) - log(rho[nb.dn])) / (nb.Pup - nb.Pdn);
        KT_err[j] = fabs(rho_err[nb.dn] / rho[nb.dn] + rho_err[nb.up] / rho[nb.up]) / (nb.Pup - nb.Pdn);
      }
    }
  }
  // truncation term (error_compresibilidad_log, KTderivate.c:483-564); it
  // reads K_T only, so the order of the points does not matter
  for (int j = 0; j < count; ++j) {
    const Neighbours& nb = nbs[j];
    if (KT[j] == NV || KT_err[j] == NV) continue;
    if (nb.up2 != -1 && nb.dn2 != -1) {
      if (rho[nb.up2] < gas_rho || rho[nb.up] < gas_rho || rho[nb.dn2] < gas_rho || rho[nb.dn] < gas_rho)
        continue;
      const double dP2 = (nb.Pup2 - nb.Pdn2) / 4.0;
      const double errDer =
          fabs(0.5 * log(rho[nb.up2]) - log(rho[nb.up]) + log(rho[nb.dn]) - 0.5 * log(rho[nb.dn2])) /
          (6.0 * dP2);
      KT_err[j] = sqrt(KT_err[j] * KT_err[j] + errDer * errDer);
    } else if (nb.up != -1 && nb.dn != -1) {
      if (KT[nb.up] == NV || KT[nb.dn] == NV) continue;
      const double errDer = fabs(KT[nb.dn] - 2 * KT[j] + KT[nb.up]) / 6.0;
      KT_err[j] = sqrt(KT_err[j] * KT_err[j] + errDer * errDer);
    }
  }
  return CVF_OK;
}

}  // extern "C"
