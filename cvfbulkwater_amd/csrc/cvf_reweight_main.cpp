// cvf_reweight: drop-in for the reference's histogram reweighting program
// (HistogramReweighting/histogram_reweighting_mixed_paramSI_norm_x.f90, run
// by histogram_reweighting_mixed_paramSI_norm_x.sh).  Run in (or give) a
// directory holding the `input` file the driver script writes and the
// data_L<L>_T<T>_P<P> files of (h, rho) samples:
//
//   cvf_reweight [dir] [--device N]
//
// Outputs, as the Fortran program's:
//   Constants_L<L>                reweighting constants (read back when
//                                 present and matching, f90:386-495)
//   Binder_L..._energy_density.dat  T P U_e U_rho per scan point
//   Binder_L..._mA.._MA...dat     T P S U_M per scan point and angle
//   Confront_L...Ising3D_{Liu,KL}.dat  T P S distance, a blank line per T
//   Histogram_Reweighted_OP_L...  the order-parameter histogram at the
//                                 liu_3D minimum with its error bars
// and on stderr the constants and the five minima (Liu 2D, Liu 3D, Zero,
// Height, Zero + Height) with their S, T, P and B.
//
// The scan's (T, P) points are reweighted as one batch on the GPU, with
// the cumulants and order-parameter histograms reduced there.  Numbers in
// the files are written with 17 significant digits.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cvf_api.h"
#include "cvf_reweight.h"

namespace {

constexpr double sl(float x) { return (double)x; }
constexpr double kTs = sl(140.5678f), kTi = sl(185.4676f), kPs = sl(4694.59f), kPi = sl(-2111.5552f);

struct Input {
  int n_sim = 0, steps = 0, L = 0;
  double Tmin, Tmax, dT, Pmin, Pmax, dP, Smin, Smax, dS, bin_factor;
  std::vector<double> T_hat, P_hat;
  std::vector<int64_t> counts;
};

// `input` as the driver script writes it (…_x.sh:167-179, 200-203): thirteen
// "KEY value" lines, then one "T P count" line per simulation
bool read_input(const std::string& path, Input& in) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return false;
  char key[64];
  double* dst[10] = {&in.Tmin, &in.Tmax, &in.dT, &in.Pmin, &in.Pmax, &in.dP, &in.Smin, &in.Smax, &in.dS,
                     &in.bin_factor};
  bool ok = fscanf(f, "%63s %d", key, &in.n_sim) == 2 && fscanf(f, "%63s %d", key, &in.steps) == 2 &&
            fscanf(f, "%63s %d", key, &in.L) == 2;
  for (int k = 0; k < 10 && ok; ++k) ok = fscanf(f, "%63s %lf", key, dst[k]) == 2;
  ok = ok && in.n_sim > 0;
  for (int i = 0; i < in.n_sim && ok; ++i) {
    double t, p;
    long long n;
    ok = fscanf(f, "%lf %lf %lld", &t, &p, &n) == 3;
    in.T_hat.push_back(t);
    in.P_hat.push_back(p);
    in.counts.push_back(n);
  }
  fclose(f);
  return ok;
}

std::string data_name(int L, double T_hat, double P_hat) {
  char b[128];
  // F9.7 / F6.4 (F6.3 below zero) fields (f90:143-155)
  snprintf(b, sizeof b, P_hat >= 0 ? "data_L%d_T%9.7f_P%6.4f" : "data_L%d_T%9.7f_P%6.3f", L, T_hat, P_hat);
  return b;
}

// pressure field of the output names: F7.1 / F6.1 / F5.1 (f90:534-625,
// 631-714); the histogram file's choice is strict at +-1000 (2293-2309)
std::string p_field(double P, bool strict) {
  char b[32];
  if (strict ? P < -1000.0 : P <= -1000.0)
    snprintf(b, sizeof b, "%7.1f", P);
  else if (P < 0.0 || (strict ? P > 1000.0 : P >= 1000.0))
    snprintf(b, sizeof b, "%6.1f", P);
  else
    snprintf(b, sizeof b, "%5.1f", P);
  return b;
}

void row(FILE* f, std::initializer_list<double> v) {
  for (double x : v) fprintf(f, " %.17g", x);
  fprintf(f, "\n");
}

FILE* open_w(const std::string& path) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  return f;
}

[[noreturn]] void die(const char* what, int code) {
  fprintf(stderr, "%s: %s\n", what, cvf_last_error());
  exit(code);
}

struct Best {
  double v = sl(1E12f), S = 0, T = 0, P = 0, B = 0;
};

}  // namespace

int main(int argc, char** argv) {
  std::string dir = ".";
  int device = 0;
  for (int a = 1; a < argc; ++a) {
    if (!strcmp(argv[a], "--device") && a + 1 < argc)
      device = atoi(argv[++a]);
    else
      dir = argv[a];
  }
  Input in;
  if (!read_input(dir + "/input", in)) {
    fprintf(stderr, "cannot read %s/input\n", dir.c_str());
    return 2;
  }
  if (in.Tmax <= in.Tmin) return fprintf(stderr, " Error: Tmax <= Tmin\n"), 1;
  if (in.Pmax <= in.Pmin) return fprintf(stderr, " Error: Pmax <= Pmin\n"), 1;
  if (in.Smax <= in.Smin) return fprintf(stderr, " Error: Smax <= Smin\n"), 1;
  const int L = in.L, ns = in.n_sim;
  // samples, simulation after simulation
  std::vector<double> h_rho;
  for (int i = 0; i < ns; ++i) {
    const std::string path = dir + "/" + data_name(L, in.T_hat[i], in.P_hat[i]);
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return fprintf(stderr, "cannot open %s\n", path.c_str()), 2;
    for (int64_t j = 0; j < in.counts[i]; ++j) {
      double h, r;
      if (fscanf(f, "%lf %lf", &h, &r) != 2) return fprintf(stderr, "short file %s\n", path.c_str()), 2;
      h_rho.push_back(h);
      h_rho.push_back(r);
    }
    fclose(f);
  }
  cvf_reweight* rw = nullptr;
  if (cvf_rw_create(L, ns, in.T_hat.data(), in.P_hat.data(), in.counts.data(), h_rho.data(), in.bin_factor, device,
                    &rw) != 0)
    die("cvf_rw_create", 2);
  cvf_rw_info info;
  cvf_rw_get_info(rw, &info);
  std::vector<double> T(ns), P(ns), ctes(ns), prev(ns);
  cvf_rw_sim_points(rw, T.data(), P.data());
  // constants: the file when it matches this input, else solved
  const std::string cpath = dir + "/Constants_L" + std::to_string(L);
  bool read_ok = false, exists = false;
  if (FILE* f = fopen(cpath.c_str(), "r")) {
    exists = read_ok = true;
    int i = 0;
    double tc, pc, vc;
    while (fscanf(f, "%lf %lf %lf", &tc, &pc, &vc) == 3) {
      if (i >= ns) {
        read_ok = false;
        break;
      }
      ctes[i] = vc;
      if (fabs(tc - in.T_hat[i]) > sl(1E-10f) || fabs(pc - in.P_hat[i]) > sl(1E-10f)) read_ok = false;
      ++i;
    }
    if (i != ns) read_ok = false;
    fclose(f);
    if (!read_ok) fprintf(stderr, " WARNING: the constants file does not match; recalculating, file kept\n");
  }
  if (read_ok) {
    cvf_rw_set_constants(rw, ctes.data());
  } else {
    const int rc = cvf_rw_solve_constants(rw, in.steps, ctes.data(), prev.data());
    if (rc == CVF_RW_NOT_CONVERGED) {
      fprintf(stderr, " %s\n", cvf_last_error());
      return 1;
    }
    if (rc != 0) die("cvf_rw_solve_constants", 2);
  }
  fprintf(stderr, " constants");
  for (double c : ctes) fprintf(stderr, " %.17g", c);
  fprintf(stderr, "\n");
  if (!exists) {
    FILE* f = open_w(cpath);
    for (int i = 0; i < ns; ++i) row(f, {in.T_hat[i], in.P_hat[i], ctes[i]});
    fclose(f);
  }
  // output files (f90:534-714)
  const double Tmin_SI = kTs * in.Tmin + kTi, Tmax_SI = kTs * in.Tmax + kTi;
  char tag[256], tagS[256];
  const std::string pf = p_field(P[0], false);
  snprintf(tag, sizeof tag, "_L%d_P%sbar_Tm%9.5fK_TM%9.5fK", L, pf.c_str(), Tmin_SI, Tmax_SI);
  snprintf(tagS, sizeof tagS, "%s_mA%8.6f_MA%8.6f", tag, in.Smin, in.Smax);
  FILE* f20 = open_w(dir + "/Binder" + tagS + ".dat");
  FILE* f21 = open_w(dir + "/Binder" + tag + "_energy_density.dat");
  FILE* f12 = open_w(dir + "/Confront" + tagS + "Ising3D_Liu.dat");
  FILE* f13 = open_w(dir + "/Confront" + tagS + "Ising3D_KL.dat");
  // scan points, accumulated as the Fortran loops do (f90:928-1001)
  std::vector<double> TK, PB;
  for (double P1 = in.Pmin; P1 <= in.Pmax; P1 = P1 + in.dP)
    for (double T1 = in.Tmin; T1 <= in.Tmax; T1 = T1 + in.dT) {
      TK.push_back(T1 * kTs + kTi);
      PB.push_back(P1 * kPs + kPi);
    }
  const int nm = info.n_bin_mixed;
  std::vector<double> x(nm), p(nm);
  Best best[5];  // liu_2D, liu_3D, zero, height, zero + height
  double alpha = in.Smin;
  // the angles of every scan point (f90:990-995)
  std::vector<double> angles;
  for (double S = in.Smin; S < in.Smax; S = S + in.dS) angles.push_back(S);
  const int na = (int)angles.size();
  if (na > 4096) return fprintf(stderr, "more than 4096 angles\n"), 2;
  // targets per pass: at most 4096 (target, angle) pairs and about 2 GiB of
  // densities, reduced on the device
  const int64_t chunk = std::max<int64_t>(
      1, std::min<int64_t>(4096 / na, (int64_t(1) << 28) / info.n_occupied));
  std::vector<double> ued, xs, ps, sts, scs;
  for (size_t t0 = 0; t0 < TK.size(); t0 += chunk) {
    const int nt = (int)std::min<size_t>(chunk, TK.size() - t0);
    ued.resize(2 * (size_t)nt);
    xs.resize((size_t)nt * na * nm);
    ps.resize(xs.size());
    sts.resize((size_t)nt * na * 3);
    scs.resize((size_t)nt * na * 6);
    if (cvf_rw_scan(rw, nt, TK.data() + t0, PB.data() + t0, na, angles.data(), ued.data(), xs.data(), ps.data(),
                    sts.data(), scs.data()) != 0)
      die("scan", 255);
    for (int t = 0; t < nt; ++t) {
      const double Tk = TK[t0 + t], Pb = PB[t0 + t];
      row(f21, {Tk, Pb, ued[2 * t], ued[2 * t + 1]});
      for (int a = 0; a < na; ++a) {
        const size_t ta = (size_t)t * na + a;
        const double S = angles[a];
        alpha = S;
        row(f20, {Tk, Pb, S, sts[3 * ta + 2]});
        // the reference's scale loop over B runs once, at B = 1 (f90:800-852)
        for (double B = 1; B < sl(1.001f); B = B + sl(0.01f)) {
          std::vector<double> xb(nm);
          for (int i = 0; i < nm; ++i) xb[i] = B * xs[ta * nm + i];
          double sc[6];
          cvf_rw_ising_scores(xb.data(), ps.data() + ta * nm, nm, sc);
          row(f12, {Tk, Pb, S, sc[3]});
          row(f13, {Tk, Pb, S, sc[1]});
          const double vals[5] = {sc[2], sc[3], sc[4], sc[5], sc[4] + sc[5]};
          for (int k = 0; k < 5; ++k)
            if (best[k].v > vals[k]) best[k] = Best{vals[k], alpha, Tk, Pb, B};
        }
      }
      fprintf(f12, "\n");
      fprintf(f13, "\n");
    }
  }
  fclose(f20);
  fclose(f21);
  fclose(f12);
  fclose(f13);
  static const char* what[5] = {"Liu 2D=", "Liu 3D=", "Zero Test=", "Height Test=", "Zero + Height Test="};
  for (int k = 0; k < 5; ++k) {
    // (A8,F9.6,A3,F9.5,A4,F6.1) with format reversion for the last pair
    fprintf(stderr, " %s %.17g\n", what[k], best[k].v);
    fprintf(stderr, "%8s%9.6f%3s%9.5f%4s%6.1f\n%8s%9.6f\n", "S=", best[k].S, " T=", best[k].T, "K P=", best[k].P,
            "bar B=", best[k].B);
  }
  // error bars from the simulation nearest to the liu_3D minimum (f90:931-965)
  const Best& b3 = best[1];
  int err_k = 0;
  double dmin = 0;
  for (int k = 0; k < ns; ++k) {
    const double d = (T[k] - b3.T) * (T[k] - b3.T) + (P[k] - b3.P) * (P[k] - b3.P);
    if (k == 0 || d < dmin) dmin = d, err_k = k;
  }
  std::vector<double> eh(nm), ex(nm);
  if (cvf_rw_reweight(rw, 1, &T[err_k], &P[err_k], nullptr) != 0) die("cvf_rw_reweight", 2);
  if (cvf_rw_errors(rw, 0, err_k, b3.S, eh.data(), ex.data()) != 0) die("error bars", 255);
  // Print_Histogram (f90:2256-2323): the histogram at the liu_3D minimum,
  // projected at the last angle of the scan (the program's global alpha)
  // and named after the minimum's angle
  if (cvf_rw_reweight(rw, 1, &b3.T, &b3.P, nullptr) != 0) die("cvf_rw_reweight", 2);
  double st[3];
  if (cvf_rw_order_parameter(rw, 0, alpha, x.data(), p.data(), st) != 0) die("order parameter", 255);
  char name[512];
  snprintf(name, sizeof name, "/Histogram_Reweighted_OP_L%d_T%9.5f_P%s_A%8.6f_B%5.3f", L, b3.T,
           p_field(b3.P, true).c_str(), b3.S, b3.B);
  FILE* fh = open_w(dir + name);
  fprintf(fh, " #Data simulations (Temperature internal units): ");
  for (double t : in.T_hat) fprintf(fh, " %.17g", t);
  fprintf(fh, "\n");
  for (int i = 0; i < nm; ++i) row(fh, {x[i] * b3.B, p[i], ex[i], eh[i]});
  fclose(fh);
  printf(" alpha = %.17g\n Mixed Param (min) = %.17g ; Mixed_param (max) = %.17g\n Normalization factor= %.17g\n",
         b3.S, x[0], x[nm - 1], st[1]);
  fprintf(stderr, " normalization %.17g\n", st[1]);
  cvf_rw_destroy(rw);
  return 0;
}
