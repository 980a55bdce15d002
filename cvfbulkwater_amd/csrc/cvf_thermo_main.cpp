// cvf_thermo: drop-in for the reference's CVF_ThermodynamicAnalysis pipeline.
// Run in a directory with one L<L>_T<T>_P<P>/ folder per state point, each
// holding its data.out.  Three stages, each reading the previous stage's
// text files as the scripts do (`cvf_thermo [stage] [dir]`, stage = all by
// default):
//
//   averages  script_analisi.sh + filter_isobars.sh:
//             L<L>_average_data.dat (16 columns per state point) and
//             L<L>_P<P>.dat (the rows of one isobar, sorted by T)
//   isobars   calculadora.sh (+ AlphaPderivate, CPderivate):
//             L<L>_P<P>_{energy,enthalpy,volume,density,nhb,ncoop,rcell,
//             Cpfluc,KT,AlphaP,CP}.dat
//   convert   conversor.sh (+ AlphaPderivate, KTderivate): the .dat.SI files,
//             L<L>_P<P>_{AlphaPderivate,energy}.dat.SI and
//             L<L>_P<P in MPa>_KTderivate.dat.SI
//
// Numbers go between the stages as text printed the way awk prints them
// (integers as integers, otherwise "%.6g"), and the derivative programs'
// outputs as "%lf" / "%e", so every output is the same function of the
// data.out files as the scripts'.  Isobars are those present in the
// directory (the scripts carry a fixed list of pressures).
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "cvf_thermo.h"

namespace {

// SI conversion, conversor.sh:2-13 (slope, intercept)
constexpr double kTs = 140.5678, kTi = 185.4676;   // T -> K
constexpr double kPs = 469.459, kPi = -211.15552;   // P -> MPa
constexpr double kRhos = 1527.3, kRhoi = -23.1017;  // density -> kg/m^3
constexpr double kCPs = 4.455, kCPi = 3.56843;      // C_P -> J/(g mol)
constexpr double kHs = 11.272132, kHi = 9.028924;   // enthalpy -> kJ/mol
constexpr double kGasRho = 500.0;                   // KTderivate.c:8, kg/m^3

using Row = std::vector<std::string>;
using Table = std::vector<Row>;

// awk's output of a number: integers as integers (through a C integer, so
// -0 prints as 0), otherwise OFMT "%.6g"
std::string num(double v) {
  char b[64];
  if (std::isfinite(v) && v == floor(v) && fabs(v) < 1e15)
    snprintf(b, sizeof b, "%lld", (long long)v);
  else
    snprintf(b, sizeof b, "%.6g", v);
  return b;
}
std::string fmt(const char* f, double v) {
  char b[64];
  snprintf(b, sizeof b, f, v);
  return b;
}
// awk's numeric value of a field (empty -> 0)
double val(const Row& r, size_t k) { return k < r.size() ? strtod(r[k].c_str(), nullptr) : 0.0; }

[[noreturn]] void die(const std::string& m) {
  fprintf(stderr, "%s\n", m.c_str());
  exit(1);
}

Table read_table(const std::string& path) {
  Table t;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return t;
  char line[4096];
  while (fgets(line, sizeof line, f)) {
    Row r;
    for (char* tok = strtok(line, " \t\r\n"); tok; tok = strtok(nullptr, " \t\r\n")) r.push_back(tok);
    if (!r.empty()) t.push_back(r);
  }
  fclose(f);
  return t;
}

void write_table(const std::string& path, const Table& t, const char* sep = " ") {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) die("Can't create " + path);
  for (const auto& r : t) {
    for (size_t k = 0; k < r.size(); ++k) fprintf(f, "%s%s", k ? sep : "", r[k].c_str());
    fprintf(f, "\n");
  }
  fclose(f);
}

// sort -g on the first field (stable for equal keys)
void sort_g(Table& t) {
  std::stable_sort(t.begin(), t.end(), [](const Row& a, const Row& b) { return val(a, 0) < val(b, 0); });
}

std::vector<std::string> list_dir() {
  std::vector<std::string> v;
  if (DIR* d = opendir(".")) {
    while (dirent* e = readdir(d)) v.push_back(e->d_name);
    closedir(d);
  }
  std::sort(v.begin(), v.end());
  return v;
}
bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}
bool is_dir(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// numbers of a folder name, as grep -o -E '[0-9.-]+' finds them
std::vector<std::string> numbers_in(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if ((c >= '0' && c <= '9') || c == '.' || c == '-') {
      cur += c;
    } else if (!cur.empty()) {
      out.push_back(cur);
      cur.clear();
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// L<L>_P<P><suffix> -> (L, P); false for other names
bool isobar_name(const std::string& n, const std::string& suffix, std::string* L, std::string* P) {
  if (n.size() < 4 || n[0] != 'L' || !ends_with(n, suffix)) return false;
  const size_t u = n.find("_P");
  if (u == std::string::npos || u < 2) return false;
  const std::string mid = n.substr(u + 2, n.size() - suffix.size() - u - 2);
  if (mid.empty() || mid.find('_') != std::string::npos) return false;
  *L = n.substr(1, u - 1);
  *P = mid;
  return true;
}

// AlphaPderivate / CPderivate on rows (T, X, var, n): the "%lf %lf %lf" rows
Table derivative(const Table& in, int mode) {
  std::vector<double> T, X, V;
  std::vector<int32_t> n;
  for (const auto& r : in) {
    T.push_back(val(r, 0));
    X.push_back(val(r, 1));
    V.push_back(val(r, 2));
    n.push_back((int32_t)val(r, 3));
  }
  const int c = (int)T.size();
  std::vector<double> to(c + 1), d(c + 1), e(c + 1);
  const int m = cvf_thermo_derivative(T.data(), X.data(), V.data(), n.data(), c, mode, to.data(), d.data(),
                                      e.data());
  Table out;
  for (int i = 0; i < m; ++i) out.push_back({fmt("%lf", to[i]), fmt("%lf", d[i]), fmt("%lf", e[i])});
  return out;
}

// ---------------------------------------------------------------- stages
// script_analisi.sh + filter_isobars.sh
void stage_averages() {
  std::map<std::string, Table> by_L;
  for (const auto& d : list_dir()) {
    if (d.empty() || d[0] != 'L' || d.find("_T") == std::string::npos || d.find("_P") == std::string::npos ||
        !is_dir(d))
      continue;
    const auto nums = numbers_in(d);
    if (nums.size() < 3) continue;
    std::vector<double> rows;
    for (const auto& r : read_table(d + "/data.out")) {
      if (r.size() < 7) continue;
      for (int k = 1; k <= 6; ++k) rows.push_back(val(r, k));  // e, h, v, N_HB, m, r_cell
    }
    const int64_t nr = (int64_t)rows.size() / 6;
    if (nr == 0) {
      fprintf(stderr, "warning: %s/data.out has no rows\n", d.c_str());
      continue;
    }
    double a[12];
    cvf_thermo_averages(rows.data(), nr, 6, a);
    Row out = {nums[0], nums[2], nums[1]};  // L, P, T
    for (double v : a) out.push_back(num(v));
    out.push_back(num((double)nr));
    by_L[nums[0]].push_back(out);
  }
  for (auto& [L, t] : by_L) {
    write_table("L" + L + "_average_data.dat", t, "  ");
    std::map<std::string, Table> iso;
    for (const auto& r : t) iso[r[1]].push_back(r);
    for (auto& [P, rows] : iso) {
      std::stable_sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return val(a, 2) < val(b, 2); });
      write_table("L" + L + "_P" + P + ".dat", rows, "  ");
    }
  }
}

// calculadora.sh
void stage_isobars() {
  for (const auto& f : list_dir()) {
    std::string L, P;
    if (!isobar_name(f, ".dat", &L, &P)) continue;
    const Table t = read_table(f);
    const std::string base = f.substr(0, f.size() - 4);
    Table en, hn, vo, de, nh, nc, rc, cp, kt, ai, ci;
    for (const auto& r : t) {
      if (r.size() < 16) continue;
      const double l = val(r, 0), T = val(r, 2), n = val(r, 15);
      const double N = l * l * l;
      auto c = [&](int k) { return val(r, k - 1); };  // awk's $k
      en.push_back({r[2], num(c(4) / N), num(sqrt(3 * c(5) / n) / N)});
      hn.push_back({r[2], num(c(6) / N), num(sqrt(3 * c(7) / n) / N)});
      vo.push_back({r[2], num(c(8) / N), num(sqrt(3 * c(9) / n) / N)});
      de.push_back({r[2], num(N / c(8)), num(N * sqrt(3 * c(9) / n) / (c(8) * c(8)))});
      nh.push_back({r[2], num(c(10) / (2 * l * l * l)), num(sqrt(3 * c(11) / n) / (2 * l * l * l))});
      nc.push_back({r[2], r[11], num(sqrt(3 * c(13) / n))});
      rc.push_back({r[2], r[13], num(sqrt(c(15) / n))});
      cp.push_back({r[2], num(c(7) / (T * T * l * l * l))});
      kt.push_back({r[2], num(c(9) / (T * c(8)))});
      ai.push_back({r[2], r[7], r[8], r[15]});
      ci.push_back({r[2], num(c(6) / (l * l * l)), num(c(7) / (l * l * l * l * l * l)), r[15]});
    }
    for (Table* x : {&en, &de, &nh, &nc, &rc, &cp}) sort_g(*x);
    write_table(base + "_energy.dat", en);
    write_table(base + "_enthalpy.dat", hn);
    write_table(base + "_volume.dat", vo);
    write_table(base + "_density.dat", de);
    write_table(base + "_nhb.dat", nh);
    write_table(base + "_ncoop.dat", nc);
    write_table(base + "_rcell.dat", rc);
    write_table(base + "_Cpfluc.dat", cp);
    write_table(base + "_KT.dat", kt);
    write_table(base + "_AlphaP.dat", derivative(ai, 1));
    write_table(base + "_CP.dat", derivative(ci, 0));
  }
}

// conversor.sh
void stage_convert() {
  auto each = [](const std::string& suffix, auto&& fn) {
    for (const auto& f : list_dir()) {
      std::string L, P;
      if (isobar_name(f, suffix, &L, &P)) fn(f, L, P);
    }
  };
  auto tsi = [](const Row& r) { return num(val(r, 0) * kTs + kTi); };
  each("_density.dat", [&](const std::string& f, const std::string&, const std::string&) {
    Table o;
    for (const auto& r : read_table(f)) o.push_back({tsi(r), num(val(r, 1) * kRhos + kRhoi), num(val(r, 2) * kRhos)});
    write_table(f + ".SI", o);
  });
  each("_enthalpy.dat", [&](const std::string& f, const std::string&, const std::string&) {
    Table o;
    // the error column is scaled by Hi, as the script has it
    for (const auto& r : read_table(f))
      o.push_back({tsi(r), num(val(r, 1) * kHs + kHi * val(r, 0)), num(val(r, 2) * kHi)});
    write_table(f + ".SI", o);
  });
  for (const char* sfx : {"_nhb.dat", "_ncoop.dat"})
    each(sfx, [&](const std::string& f, const std::string&, const std::string&) {
      Table o;
      for (const auto& r : read_table(f)) {
        Row x = {tsi(r)};
        for (size_t k = 1; k < 3; ++k) x.push_back(k < r.size() ? r[k] : "");
        o.push_back(x);
      }
      write_table(f + ".SI", o);
    });
  each("_Cpfluc.dat", [&](const std::string& f, const std::string&, const std::string&) {
    Table o;
    for (const auto& r : read_table(f)) o.push_back({tsi(r), num(val(r, 1) * kCPs + kCPi), num(val(r, 2) * kCPs)});
    write_table(f + ".SI", o);
  });
  // alpha_P = (1/V) dV/dT from the SI density (conversor.sh:64-76), and
  // E = H - P V in SI (conversor.sh:80-88)
  each("_density.dat.SI", [&](const std::string& f, const std::string& L, const std::string& P) {
    const std::string base = "L" + L + "_P" + P;
    const Table d = read_table(f), iso = read_table(base + ".dat");
    Table in;
    for (size_t i = 0; i < d.size(); ++i) {
      const double rho = val(d[i], 1), e = val(d[i], 2);
      const std::string n = i < iso.size() && iso[i].size() > 15 ? iso[i][15] : "";
      // specific-volume variance, then x n, each printed by awk
      const double v2 = strtod(num(e * e / (3 * rho * rho * rho * rho)).c_str(), nullptr);
      in.push_back({d[i][0], num(1.0 / rho), num(v2 * strtod(n.c_str(), nullptr)), n.empty() ? "0" : n});
    }
    Table o;
    for (const auto& r : derivative(in, 1)) {
      if (r[1] == "-nan" || r[1] == "inf" || r[1] == "-inf" || r[2] == "inf") continue;
      o.push_back({r[0], num(val(r, 1) * 1000), num(val(r, 2) * 1000)});
    }
    write_table(base + "_AlphaPderivate.dat.SI", o);
    const Table h = read_table(base + "_enthalpy.dat.SI");
    if (!h.empty()) {
      const double PSI = strtod(P.c_str(), nullptr) * kPs + kPi;
      Table en;
      for (size_t i = 0; i < h.size() && i < d.size(); ++i) {
        const double rho = val(d[i], 1), e = val(d[i], 2);
        en.push_back({h[i][0], num(val(h[i], 1) - 18 * PSI / rho), num(val(h[i], 2) - 18 * PSI * e / (rho * rho))});
      }
      write_table(base + "_energy.dat.SI", en);
    }
  });
  // K_T along isotherms from the averages table, written per isobar
  // (conversor.sh:95-128, KTderivate.c)
  for (const auto& f : list_dir()) {
    if (f.size() < 5 || f[0] != 'L' || !ends_with(f, "_average_data.dat")) continue;
    const std::string L = f.substr(1, f.size() - 1 - strlen("_average_data.dat"));
    Table in;
    std::vector<std::string> isobars;
    for (const auto& r : read_table(f)) {
      if (r.size() < 16) continue;
      const double l = val(r, 0), N = l * l * l;
      in.push_back({num(kTs * val(r, 2) + kTi), num(kPs * val(r, 1) + kPi), num(kRhos * N / val(r, 7) + kRhoi),
                    num(kRhos * N * sqrt(3 * val(r, 8) / val(r, 15)) / (val(r, 7) * val(r, 7)))});
      if (std::find(isobars.begin(), isobars.end(), r[1]) == isobars.end()) isobars.push_back(r[1]);
    }
    const int n = (int)in.size();
    std::vector<double> T(n + 1), Pr(n + 1), rho(n + 1), er(n + 1), KT(n + 1), EKT(n + 1);
    for (int i = 0; i < n; ++i) {
      T[i] = val(in[i], 0);
      Pr[i] = val(in[i], 1);
      rho[i] = val(in[i], 2);
      er[i] = val(in[i], 3);
    }
    cvf_thermo_kt(T.data(), Pr.data(), rho.data(), er.data(), n, kGasRho, KT.data(), EKT.data());
    for (const auto& Ps : isobars) {
      const double plist = strtod(num(kPs * strtod(Ps.c_str(), nullptr) + kPi).c_str(), nullptr);
      Table o;
      for (int k = 0; k < n; ++k) {
        if (KT[k] == CVF_KT_NOT_VALUE || !(fabs(Pr[k] - plist) < 1e-5)) continue;
        const Row r = {fmt("%lf", T[k]), fmt("%lf", KT[k]), fmt("%e", EKT[k]), fmt("%lf", Pr[k])};
        o.push_back({r[0], num(val(r, 1) * 1000), num(val(r, 2) * 1000), r[3]});
      }
      sort_g(o);
      write_table("L" + L + "_P" + fmt("%lf", plist) + "_KTderivate.dat.SI", o);
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string stage = "all", dir = ".";
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    if (s == "all" || s == "averages" || s == "isobars" || s == "convert")
      stage = s;
    else
      dir = s;
  }
  if (chdir(dir.c_str()) != 0) die("Can't enter " + dir);
  if (stage == "all" || stage == "averages") stage_averages();
  if (stage == "all" || stage == "isobars") stage_isobars();
  if (stage == "all" || stage == "convert") stage_convert();
  return 0;
}
