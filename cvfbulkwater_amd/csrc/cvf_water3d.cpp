// cvf_water3d: drop-in replacement for the reference executable gpu_water3D
// (SOURCE/gpu_water3D.cu main(), 279-1321).  Same working-directory file
// contract: reads input_data, input_config (Flag_config = 1) and distances
// (Flag_distances = 1); writes data.out, cluster.out, config.out,
// water3D.log and distances (Flag_distances = 0).  The lattice size, a
// compile-time constant in the reference (35-37), is read from an optional
// 15th input_data line "Lattice_Size L" or from "-L <L>" (default 32).
// "-schedule fast" selects the engine's two-colour sigma passes instead of
// the reference's half-sweeps (same model, different trajectory).
// All simulation work goes through the C ABI of libcvf.so.
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "cvf_api.h"

namespace {

[[noreturn]] void die(const char* msg) {
  fprintf(stderr, "%s\n", msg);
  exit(1);
}

void check(int rc, const char* what) {
  if (rc < 0) {
    fprintf(stderr, "Error: %s: %s\n", what, cvf_last_error());
    exit(1);
  }
}

struct Input {
  float P = 0, T = 0, R_cutoff = 0;
  long long seed = 0;
  int n_metropolis = 0, n_cluster = 0, n_equil = 0, n_loops = 0, n_measure = 0;
  int flag_config = 0, flag_distances = 0, flag_correlation = 0, flag_chessboard = 0, device = 0;
  int L = 32;
};

// input_data, gpu_water3D.cu:292-386: 14 positional "name value" pairs.
Input read_input_data(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) die("Can't open input file");
  Input in;
  char name[256];
  auto rd_f = [&](float* v, const char* what) {
    if (fscanf(f, "%255s %f", name, v) != 2) {
      fprintf(stderr, "Error [%s]: Can't read input file\n", what);
      fclose(f);
      exit(1);
    }
  };
  auto rd_i = [&](int* v, const char* what) {
    if (fscanf(f, "%255s %d", name, v) != 2) {
      fprintf(stderr, "Error[%s]: Can't read input file\n", what);
      fclose(f);
      exit(1);
    }
  };
  rd_f(&in.P, "Pressure");
  rd_f(&in.T, "Temperature");
  if (fscanf(f, "%255s %lld", name, &in.seed) != 2) {
    fprintf(stderr, "Error[Seed]: Can't read input file\n");
    fclose(f);
    exit(1);
  }
  rd_i(&in.n_metropolis, "Metropolis_Steps");
  rd_i(&in.n_cluster, "Cluster_Steps");
  rd_i(&in.n_equil, "Equilibration_Steps");
  rd_i(&in.n_loops, "Sampling_Steps");
  rd_i(&in.n_measure, "Sampling_Interval");
  rd_f(&in.R_cutoff, "R_cutoff");
  rd_i(&in.flag_config, "Flag_config");
  rd_i(&in.flag_distances, "Flag_distances");
  rd_i(&in.flag_correlation, "Flag_correlation");
  rd_i(&in.flag_chessboard, "Flag_chessboard");
  rd_i(&in.device, "Device_ID");
  int L = 0;
  if (fscanf(f, "%255s %d", name, &L) == 2 && L > 0) in.L = L;  // optional extension
  fclose(f);
  return in;
}

// input_config, gpu_water3D.cu:730-792: N rows "s0..s5 e0..e5", arm-major.
void read_input_config(const char* path, int N, std::vector<uint8_t>& s, std::vector<uint8_t>& e) {
  FILE* f = fopen(path, "r");
  if (!f) {
    fprintf(stderr, "Can't open the initial configuration file\n");
    fprintf(stderr, "File named 'input_config' is expected\n");
    exit(1);
  }
  s.assign(6 * (size_t)N, 0);
  e.assign(6 * (size_t)N, 0);
  int v[12];
  int i = 0;
  while (fscanf(f, "%d %d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                &v[6], &v[7], &v[8], &v[9], &v[10], &v[11]) == 12) {
    if (i >= N) {
      fprintf(stderr, "Initial configuration file error\n");
      fprintf(stderr, "There are more rows than cells in the file\n");
      fclose(f);
      exit(1);
    }
    for (int a = 0; a < 6; ++a) {
      if (v[a] < 0 || v[a] > 5) {
        fprintf(stderr, "Initial configuration file error\n");
        fprintf(stderr, "Unexpected value for spin (%d) appeared in line %d\n", v[a], i);
        fclose(f);
        exit(1);
      }
      if (v[6 + a] < 0 || v[6 + a] > 1) {
        fprintf(stderr, "Initial configuration file error\n");
        fprintf(stderr, "Unexpected value for active_bond (%d) appeared in line %d\n", v[6 + a], i);
        fclose(f);
        exit(1);
      }
      s[a * (size_t)N + i] = (uint8_t)v[a];
      e[a * (size_t)N + i] = (uint8_t)v[6 + a];
    }
    ++i;
  }
  fclose(f);
  if (i < N) {
    fprintf(stderr, "Initial configuration file error\n");
    fprintf(stderr, "Unexpected ending. Can't fill the spin array\n");
    exit(1);
  }
}

// distances file, gpu_water3D.cu:917-985.
void write_distances(const char* path, int L, const std::vector<float>& d, const std::vector<int32_t>& f) {
  FILE* o = fopen(path, "w");
  if (!o) die("Can't create distances file");
  fprintf(o, "%d %d %d\n", L, L, L);
  for (size_t k = 0; k < d.size(); ++k) fprintf(o, "%f %d\n", d[k], f[k]);
  fclose(o);
}

void read_distances(const char* path, int L, std::vector<float>& d, std::vector<int32_t>& f) {
  FILE* in = fopen(path, "r");
  if (!in) die("Can't open distances file");
  int lx, ly, lz;
  if (fscanf(in, "%d %d %d", &lx, &ly, &lz) != 3) {
    fclose(in);
    die("Error: Can't read distances file");
  }
  if (lx != L) die("Error: Distances file. X rank mismatch");
  if (ly != L) die("Error: Distances file. Y rank mismatch");
  if (lz != L) die("Error: Distances file. Z rank mismatch");
  d.clear();
  f.clear();
  float dv;
  int fv;
  while (fscanf(in, "%f %d", &dv, &fv) == 2) {
    d.push_back(dv);
    f.push_back(fv);
  }
  fclose(in);
}

}  // namespace

int main(int argc, char** argv) {
  Input in = read_input_data("input_data");
  int schedule = CVF_SCHED_REFERENCE;
  for (int a = 1; a + 1 < argc; ++a) {
    if (!strcmp(argv[a], "-L")) in.L = atoi(argv[a + 1]);
    if (!strcmp(argv[a], "-schedule")) {
      if (!strcmp(argv[a + 1], "fast"))
        schedule = CVF_SCHED_FAST;
      else if (strcmp(argv[a + 1], "reference"))
        die("Error: -schedule takes reference or fast");
    }
  }
  const int L = in.L, N = L * L * L;

  FILE* odata = fopen("data.out", "w");
  if (!odata) die("Can't create output file");
  FILE* log = fopen("water3D.log", "w");
  if (!log) die("Can't create log file");
  fprintf(log, "++++ Welcome to water3D log file ++++\n");

  cvf_params p;
  cvf_params_default(&p);
  p.L = L;
  p.P = in.P;
  p.T = in.T;
  p.seed = (uint64_t)in.seed;
  p.n_metropolis = in.n_metropolis;
  p.n_cluster = in.n_cluster;
  p.r_cutoff = in.R_cutoff;
  p.flag_config = in.flag_config;
  p.flag_chessboard = in.flag_chessboard;
  p.device_id = in.device;
  p.precision = CVF_FP64;
  p.volume_mode = CVF_VOLUME_HOST;
  p.schedule = schedule;

  // header, gpu_water3D.cu:474-509 (no random-number arrays are stored)
  fprintf(log, "RNG allocated memory: %d MB\n", 0);
  fprintf(log, " ++++++++++++++++++++++++++++++++++++++++++\n");
  fprintf(log, "Source file: %s\n", "cvf_water3d.cpp");
  fprintf(log, " 3D water model, %d x %d x %d lattice, N = %d, P = %1.2f, T = %1.2f\n", L, L, L, N, in.P,
          in.T);
  fprintf(log, " J_hb = %1.2f, J_sigma = %1.2f, v_HB = %1.2f\n", p.J_hb, p.J_sig, p.v_hb);
  fprintf(log, " Flag configuration = %d\n", in.flag_config);
  fprintf(log, " Monte Carlo Step: %d Metropolis + %d Cluster\n", in.n_metropolis, in.n_cluster);
  fprintf(log, " Metropolis: Parallel GPU algorithm\n");
  fprintf(log, " Cluster: Swendsen-Wang Parallel GPU algorithm\n");
  fprintf(log, " Device = %d\n", in.device);
  fprintf(log, " %d equilibration, %d updates/measurement, %d updates\n", in.n_equil, in.n_measure,
          in.n_loops);
  fprintf(log, " Output file %s\n", "data.out");
  fprintf(log, " Random seed %lld\n", in.seed);
  fflush(log);

  cvf_sim* sim = nullptr;
  if (in.flag_config < 0 || in.flag_config > 2) {
    fprintf(stderr, "Error: Wrong value for flag_config = %d\n", in.flag_config);
    fprintf(stderr, "Flag_config=0  Initial spins completely disordered\n");
    fprintf(stderr, "Flag_config=1  Initial spins read from file\n");
    fprintf(stderr, "Flag_config=2  Initial spins completely ordered (T=0)\n");
    exit(1);
  }
  if (in.flag_chessboard == 1 && L % 4 != 0) {
    fprintf(stderr, "Error: can't build chessboard.\n\tUse NX muliply of 4.\n");
    exit(1);
  }
  if (cvf_create(&p, &sim) < 0) {
    // set-up errors carry the reference's own message (e.g. a Device_ID
    // beyond the device count, gpu_water3D.cu:461-465, which exits with 1)
    fprintf(stderr, "%s\n", cvf_last_error());
    exit(1);
  }
  if (in.flag_chessboard == 1) fprintf(log, "Starting Chessboard\n");
  if (in.flag_correlation == 1) {
    // every step is recorded (a histogram after each): the resident kernel
    // would run one launch per step, so the fused slab plan runs instead
    // (L = 32: 36.5 against 40.2 us per step)
    cvf_tuning t;
    check(cvf_get_tuning(sim, &t), "tuning");
    if (t.fused == 2) {
      t.fused = 1;
      t.slab_planes = 0;
      t.block_chunk = 512;
      t.graph_steps = 16;
      check(cvf_set_tuning(sim, &t), "tuning");
    }
  }

  if (in.flag_config == 1) {
    std::vector<uint8_t> s, e;
    read_input_config("input_config", N, s, e);
    check(cvf_set_config(sim, 0, s.data(), e.data()), "input_config");
  }

  std::vector<float> dist(256);
  std::vector<int32_t> freq(256);
  int32_t ns = 0;
  check(cvf_get_shells(sim, 0, dist.data(), freq.data(), 256, &ns), "shells");
  dist.resize(ns);
  freq.resize(ns);
  if (in.flag_distances == 0) {
    write_distances("distances", L, dist, freq);
  } else {
    read_distances("distances", L, dist, freq);
    check(cvf_set_shells(sim, 0, dist.data(), freq.data(), (int32_t)dist.size()), "distances");
  }

  FILE* logsize = fopen("cluster.out", "w");
  if (!logsize) die("Can't open input file");

  // equilibration (gpu_water3D.cu:1176-1194), in chunks
  for (int done = 0; done < in.n_equil;) {
    const int chunk = std::min(in.n_equil - done, 1000);
    check(cvf_step(sim, chunk), "cvf_step");
    done += chunk;
  }
  // production (1204-1275): sample after step i when i % interval == 0;
  // with Flag_correlation = 1 the fraction of sigma in the most populated
  // state feeds the correlator after every step (1246-1274)
  double E = 0, E2 = 0, M = 0, M2 = 0;
  int64_t pending = 0;
  const int interval = in.n_measure > 0 ? in.n_measure : 1;
  cvf_correlator* corr = in.flag_correlation == 1 ? cvf_corr_create() : nullptr;
  std::vector<int64_t> hist;
  for (int i = 0; i < in.n_loops; ++i) {
    ++pending;
    if (corr && (i % interval == 0 || i == in.n_loops - 1)) {
      // the steps since the last sample, each followed by its histogram
      hist.resize(6 * (size_t)pending);
      check(cvf_step_histograms(sim, 0, pending, hist.data()), "sigma histograms");
      for (int64_t j = 0; j < pending; ++j) {
        double mx = -2e10;
        for (int k = 0; k < 6; ++k)
          if (hist[6 * j + k] > mx) mx = (double)hist[6 * j + k];
        cvf_corr_add(corr, mx / (double)(6 * (int64_t)N));
      }
      pending = 0;
    }
    if (i % interval == 0) {
      check(cvf_step(sim, pending), "cvf_step");
      pending = 0;
      cvf_cstat cs;
      check(cvf_cluster_stats(sim, 0, &cs), "cluster statistics");
      fprintf(logsize, "%d %d %d %d %f %f %d\n", i, cs.largest, cs.second, cs.n_perc, cs.mean, cs.var,
              cs.n_clusters);
      cvf_obs o;
      check(cvf_observe(sim, 0, &o), "measure");
      fprintf(odata, "%d %1.7e %1.7e %1.7e %1.7e %1.7e %1.7e\n", i, o.E, o.H, o.V, o.n_hb, o.m, o.r_cell);
      E += o.E;
      E2 += o.E * o.E;
      M += o.m;
      M2 += o.m * o.m;
    }
  }
  if (pending) check(cvf_step(sim, pending), "cvf_step");
  fclose(odata);
  fclose(logsize);

  // final averages (1277-1293)
  M /= in.n_loops * 1. / in.n_measure;
  M2 /= in.n_loops * 1. / in.n_measure;
  E /= in.n_loops * 1. / in.n_measure;
  E2 /= in.n_loops * 1. / in.n_measure;
  fprintf(log, "Magnetization:\t%1.6lf\t+/-\t%1.6lf\n", M, sqrt(M2 - M * M));
  fprintf(log, "Energy:\t\t%1.6lf\t+/-\t%1.6lf\n", E, sqrt(E2 - E * E));

  // config.out (1296-1305)
  std::vector<uint8_t> s(6 * (size_t)N), e(6 * (size_t)N);
  check(cvf_get_config(sim, 0, s.data(), e.data()), "config");
  FILE* oc = fopen("config.out", "w");
  if (!oc) die("Can't create config.out");
  for (int i = 0; i < N; ++i) {
    fprintf(oc, "%d %d %d %d %d %d ", s[0 * (size_t)N + i], s[1 * (size_t)N + i], s[2 * (size_t)N + i],
            s[3 * (size_t)N + i], s[4 * (size_t)N + i], s[5 * (size_t)N + i]);
    fprintf(oc, "%d %d %d %d %d %d\n", e[0 * (size_t)N + i], e[1 * (size_t)N + i], e[2 * (size_t)N + i],
            e[3 * (size_t)N + i], e[4 * (size_t)N + i], e[5 * (size_t)N + i]);
  }
  fclose(oc);
  // correlation.dat (evaluate + printCorr, 3080-3128)
  if (corr) {
    std::vector<double> t(1024), f(1024);
    const int n = cvf_corr_result(corr, t.data(), f.data(), 1024);
    FILE* oc2 = fopen("correlation.dat", "w");
    if (!oc2) die("Can't create correlation.dat");
    for (int k = 0; k < n; ++k) fprintf(oc2, "%.0f\t%f\n", t[k], f[k]);
    fclose(oc2);
    cvf_corr_destroy(corr);
  }
  // the reference logs its label-propagation scan count; union-find has none
  fprintf(log, "Max_SW_scans = %d\n", in.n_cluster > 0 ? 1 : -1);
  fprintf(log, " ** simulation done\n");
  fclose(log);
  cvf_destroy(sim);
  return 0;
}
