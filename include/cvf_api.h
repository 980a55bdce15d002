/*
 * cvf_api.h - C ABI of the MI355X CVF bulk-water Monte Carlo engine.
 *
 * The reference (lcoronas/CVFBulkWater, SOURCE/gpu_water3D.cu) exposes no
 * library interface: its hot path is the body of one executable, driven by
 * the files input_data / input_config / distances and writing data.out /
 * cluster.out / config.out / water3D.log.  Each entry point below names the
 * reference routine or file-format step it replaces.  The drop-in executable
 * cvf_water3d keeps the file contract and calls the engine only through this
 * header.
 *
 * Conventions: every int-returning function returns 0 (CVF_OK) on success
 * and a negative cvf_status on failure, with a message available from
 * cvf_last_error(); nothing here calls exit().  The library owns all device
 * memory.  Host buffers belong to the caller and use the reference's
 * arm-major uint8 layout (index arm*N + cell, cell = x + L*y + L*L*z).
 * A handle is not thread-safe and owns one HIP stream; cvf_step is
 * asynchronous on that stream, all other calls synchronise it.
 */
#ifndef CVF_API_H
#define CVF_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cvf_status {
  CVF_OK = 0,
  CVF_EINVAL = -1,   /* bad argument or parameter */
  CVF_ERANGE = -2,   /* configuration value out of range (input_config) */
  CVF_EDEVICE = -3,  /* HIP runtime error */
  CVF_ENOMEM = -4,
  CVF_EUNSUPPORTED = -5
} cvf_status;

enum { CVF_FP64 = 0, CVF_FP32_COMPENSATED = 1 };
enum { CVF_VOLUME_HOST = 0, CVF_VOLUME_DEVICE = 1 };
/* Order of the sigma moves in one Metropolis step.  REFERENCE: the
 * reference's 12 half-sweeps of randomly drawn (arm, parity)
 * (gpu_water3D.cu:1383-1417).  FAST (this engine's option, chessboard mode
 * only): two passes over the cells of colour (x + y + z) % 2, each cell
 * updating its six arms in order; the same single-move rule and the same
 * Boltzmann distribution, without the reference's trajectory. */
enum { CVF_SCHED_REFERENCE = 0, CVF_SCHED_FAST = 1 };

/* Mirrors input_data (gpu_water3D.cu:292-386) plus the reference's
 * compile-time constants (lattice size 35-37, model constants 57-64). */
typedef struct cvf_params {
  int32_t L;                 /* cubic lattice, L % 4 == 0, 4 <= L <= 252 */
  float P, T;                /* float, as gpu_water3D.cu:76 */
  uint64_t seed;
  int32_t n_metropolis;      /* Metropolis_Steps */
  int32_t n_cluster;         /* Cluster_Steps */
  float r_cutoff;            /* R_cutoff */
  int32_t flag_config;       /* 0 random, 1 caller-supplied (cvf_set_config), 2 ordered */
  int32_t flag_chessboard;   /* 1: eta cube moves; 0: eta fixed, HB gated by NHBMAX */
  int32_t device_id;
  float J_hb, J_sig, v_hb;   /* 0.5, 0.08, 0.6 */
  int32_t precision;         /* CVF_FP64 | CVF_FP32_COMPENSATED */
  int32_t volume_mode;       /* CVF_VOLUME_HOST | CVF_VOLUME_DEVICE */
  int32_t schedule;          /* CVF_SCHED_REFERENCE | CVF_SCHED_FAST */
} cvf_params;

/* One sample of measure() (gpu_water3D.cu:2014-2160). */
typedef struct cvf_obs {
  int64_t step;              /* MC steps done so far */
  double E, H, V, n_hb, m, r_cell, E_vdw;
  int64_t nhb2;              /* 2*N_HB, exact integer */
  int64_t n_sigma;           /* N_sigma, exact integer */
} cvf_obs;

/* One row of cluster.out (print_cluster_statistics, gpu_water3D.cu:2677-2721). */
typedef struct cvf_cstat {
  int32_t largest, second, n_perc, n_clusters;
  float mean, var;
} cvf_cstat;

typedef struct cvf_sim cvf_sim;

/* Defaults of SOURCE/input_data with L = 32. */
void cvf_params_default(cvf_params *p);

/* Replaces main()'s set-up (gpu_water3D.cu:449-1170): validates the
 * parameters, builds the distance-shell table (865-915), initialises sigma
 * (812-826) and the chessboard cubes (639-720) and the volume chain
 * (150, 481, 1128). */
int cvf_create(const cvf_params *p, cvf_sim **out);

/* n independent state points of uneven L advanced by one launch sequence;
 * they must agree on n_metropolis, n_cluster, precision, volume_mode,
 * schedule and device_id. */
int cvf_create_batch(const cvf_params *pts, int32_t n, cvf_sim **out);

/* input_config read (gpu_water3D.cu:730-800): sigma in [0,5], eta in {0,1};
 * cube states are re-derived from eta by the reference's decision tree
 * (3314-3371) and the eta bits rewritten from the solution table.  Returns
 * the number of cubes whose eta was not a solution (>= 0), or < 0. */
int cvf_set_config(cvf_sim *s, int32_t pt, const uint8_t *sigma6N, const uint8_t *eta6N);

/* config.out write (gpu_water3D.cu:1296-1305). */
int cvf_get_config(cvf_sim *s, int32_t pt, uint8_t *sigma6N, uint8_t *eta6N);

/* n_steps calls of MonteCarlo_Step (gpu_water3D.cu:1346-1733): eta cube
 * sweep, volume move, n_metropolis x 12 (arm, parity) half-lattice sigma
 * sweeps (or, with CVF_SCHED_FAST, n_metropolis x 2 colour passes),
 * n_cluster Swendsen-Wang steps.  Asynchronous. */
int cvf_step(cvf_sim *s, int64_t n_steps);

/* measure() (gpu_water3D.cu:2014-2160) for state point pt. */
int cvf_observe(cvf_sim *s, int32_t pt, cvf_obs *o);

/* print_cluster_statistics (gpu_water3D.cu:2677-2780) on the labels of the
 * last Swendsen-Wang step (all zero before the first, as in the reference). */
int cvf_cluster_stats(cvf_sim *s, int32_t pt, cvf_cstat *c);

/* dev_label D2H (gpu_water3D.cu:1235): canonical minimum-node labels,
 * node = arm*N + cell. */
int cvf_get_labels(cvf_sim *s, int32_t pt, int32_t *labels6N);

/* Volume-chain state (r_cell, V, E_vdW: gpu_water3D.cu:77, 150-151). */
int cvf_get_volume(cvf_sim *s, int32_t pt, double *r_cell, double *V, double *E_vdw);

/* Distance-shell table (gpu_water3D.cu:865-932; the `distances` file). */
int cvf_get_shells(cvf_sim *s, int32_t pt, float *dist, int32_t *freq, int32_t cap,
                   int32_t *nshell);

/* Replaces the Flag_distances=1 read path (gpu_water3D.cu:934-985). */
int cvf_set_shells(cvf_sim *s, int32_t pt, const float *dist, const int32_t *freq,
                   int32_t nshell);

/* Number of state points, lattice size of one, MC steps done. */
int cvf_info(cvf_sim *s, int32_t pt, int32_t *n_points, int32_t *L, int64_t *steps);

/* The handle's HIP stream (hipStream_t), for event timing by callers. */
int cvf_get_stream(cvf_sim *s, void **stream);

int cvf_sync(cvf_sim *s);
const char *cvf_last_error(void);
void cvf_destroy(cvf_sim *s);

/* Launch configuration of a handle.  It changes speed only: every choice
 * gives bit-identical trajectories (the random numbers are keyed on sites,
 * not on threads).  No reference counterpart: the reference's launch shape
 * is fixed at N/512 blocks of 512 threads (gpu_water3D.cu:68, 517-518). */
typedef struct cvf_tuning {
  int32_t streams;      /* groups of state points, each on its own stream (1..8); with
                           fused = 2, resident workgroups per CU (1 or 2) */
  int32_t block_sigma;  /* workgroup size of the sigma half-sweep (64..1024, power of 2) */
  int32_t block_eta;    /* workgroup size of the eta cube sweep */
  int32_t block_cell;   /* workgroup size of the per-cell kernels (SW, observables) */
  int32_t graph_steps;  /* MC steps per replayed hipGraph (1..64) */
  int32_t fused;        /* 1: eta sweep + 6 half-sweeps per kernel on LDS-staged z-slabs
                           where the slab fits in LDS; 0: one kernel per half-sweep;
                           2: the resident kernel (every step of a graph in one launch, the
                           lattice held in LDS, one stream group) where the points allow it
                           (reference schedule, one Metropolis sweep per step, chessboard
                           rule), else as 1.  The default is 1, or 2 with one-plane slabs
                           when every point's L exceeds 64 (no fused window fits) */
  int32_t slab_planes;  /* planes per slab of the fused kernel (<= 16), or the thickest slab
                           of the resident kernel; 0 = automatic */
  int32_t block_chunk;  /* workgroup size of the fused kernel */
  int32_t chunk_sweeps; /* half-sweeps per fused launch: 6, or 12 (a whole Metropolis pass per
                           launch where the wider window still fits in LDS); 0 = 6.  The default
                           is 12 for handles under 2^20 cells with the chessboard rule, else 6 */
} cvf_tuning;

int cvf_get_tuning(cvf_sim *s, cvf_tuning *t);

/* Launch plan of state point pt under the current tuning (no reference
 * counterpart; the reference has one launch shape): its stream group, the
 * planes per z-slab of its fused or resident launches (0 when its group
 * runs the unfused kernels), the sigma half-sweeps per fused launch of the
 * reference schedule (6 or 12; 0 when none run fused) and the kind of
 * Metropolis launch: 0 unfused kernels, 1 fused slab launches, 2 the
 * resident kernel. */
int cvf_get_plan(cvf_sim *s, int32_t pt, int32_t *group, int32_t *slab_planes, int32_t *chunk_sweeps,
                 int32_t *kind);
/* Rebuilds the groups, block tables and graphs; synchronises the handle. */
int cvf_set_tuning(cvf_sim *s, const cvf_tuning *t);

/* Project-written autotuner: times candidate tunings (fused slab kernel
 * with several slab thicknesses and workgroup sizes, the resident kernel
 * with several step-graph lengths, unfused kernels, 1/2/4 streams) for
 * max(`steps_per_trial`, the candidate's graph_steps) MC steps each, from a
 * snapshot of the state that is restored afterwards, and applies the one
 * with the shortest time per step.  The result
 * is cached in the file named by CVF_TUNE_CACHE (if set), keyed by the
 * device architecture and the batch shape.  The simulation state and step
 * count are unchanged on return. */
int cvf_autotune(cvf_sim *s, int32_t steps_per_trial, cvf_tuning *chosen);

/* Kernel classes of one MC step (profiling). */
enum {
  CVF_KCLASS_BEGIN = 0,   /* step counter, schedule, device volume move */
  CVF_KCLASS_ETA = 1,     /* cube (eta) Metropolis, per family */
  CVF_KCLASS_SIGMA = 2,   /* sigma half-lattice sweep */
  CVF_KCLASS_SW_LOCAL = 3,
  CVF_KCLASS_SW_UNION = 4,
  CVF_KCLASS_SW_FLIP = 5,
  CVF_KCLASS_CHUNK = 6,   /* fused slab kernel (eta sweep + 6 sigma half-sweeps) */
  CVF_KCLASS_RESIDENT = 7, /* resident kernel: all profiled steps in one launch (one step per
                              launch with Swendsen-Wang steps) */
  CVF_KCLASS_COUNT = 8
};

/* Instrumented cvf_step for measurement (no reference counterpart; the
 * reference's timers are commented out, gpu_water3D.cu:1180-1192): runs
 * n_steps MC steps with individual launches, each timed by start/stop HIP
 * events recorded by the dispatch itself (hipExtLaunchKernelGGL), and
 * returns the mean duration in ms of one launch of each kernel class and
 * the launch counts.  Needs the device volume mode. */
int cvf_profile_steps(cvf_sim *s, int64_t n_steps, double *kernel_ms, int64_t *launches,
                      int32_t n_classes);

/* Histogram of the 6N sigma values of point pt (the counts behind the
 * correlator observable M = max_s n_s / 6N, gpu_water3D.cu:1254-1266). */
int cvf_sigma_histogram(cvf_sim *s, int32_t pt, int64_t counts6[6]);

/* Flag_correlation's per-step record (gpu_water3D.cu:1246-1274): advances
 * n_steps MC steps (as cvf_step) and writes point pt's sigma histogram after
 * each of them, counts[6 * i .. 6 * i + 5] after step i.  The histograms
 * collect in device memory and come back in one copy per 4096 steps, so
 * the steps run back to back instead of waiting on a host round trip each. */
int cvf_step_histograms(cvf_sim *s, int32_t pt, int64_t n_steps, int64_t *counts);

/* Multi-tau correlator of the Flag_correlation = 1 mode (gpu_water3D.cu:
 * 161-191, 2983-3128; Ramirez, Sinha, Bhatt, Harmandaris, Schmid, J. Chem.
 * Phys. 133, 154103 (2010)): 32 levels of 16 points, blocking by 2.
 * cvf_corr_result returns the lag times and the normalised correlation
 * C(t)/C(0) written to correlation.dat (printCorr, 3114-3128); the return
 * value is the number of points (<= cap) or < 0. */
typedef struct cvf_correlator cvf_correlator;
cvf_correlator *cvf_corr_create(void);
void cvf_corr_add(cvf_correlator *c, double w);
int cvf_corr_result(cvf_correlator *c, double *t, double *f, int32_t cap);
void cvf_corr_destroy(cvf_correlator *c);

/* E_vdW of a shell table at r (calculate_energy_vdW, 1784-1805), evaluated
 * on the device: FP64, or FP32 terms with a compensated wavefront sum. */
int cvf_vdw_energy(int32_t device_id, int32_t L, const float *dist, const int32_t *freq,
                   int32_t nshell, float r_cutoff, const double *r, int32_t n_r,
                   int32_t precision, double *energy);

/* Pair form of E_vdW, for validating the shell table (calculate_energy_vdW,
 * gpu_water3D.cu:1784-1805, with distance_PBC, 1323-1337): the sum over all
 * minimum-image pairs of the L^3 lattice whose float distance d is below
 * R_cutoff and r*d below R_cutoff, evaluated on the device from the lattice
 * coordinates (no shell table): FP64, or FP32 terms with compensated sums.
 * Equal to the shell form up to rounding where every distance class has an
 * even count (L >= 12); at L = 8 the shell form drops half a pair of each
 * odd class (gpu_water3D.cu:912-915). */
int cvf_vdw_pair_energy(int32_t device_id, int32_t L, float r_cutoff, const double *r, int32_t n_r,
                        int32_t precision, double *energy);

/* Allocation-free form of cvf_vdw_energy: the same kernel and results, with
 * the shell table, radii and energies staged in a fixed device array that
 * each GPU holds until the process ends instead of memory allocated per
 * call.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  At most CVF_NOALLOC_MAX_SHELLS
 * shells and CVF_NOALLOC_MAX_R radii per call (CVF_EINVAL beyond).  Calls
 * share their GPU's array: calls that overlap on one GPU race. */
#define CVF_NOALLOC_MAX_SHELLS 512
#define CVF_NOALLOC_MAX_R 4096
int cvf_vdw_energy_noalloc(int32_t device_id, int32_t L, const float *dist, const int32_t *freq,
                           int32_t nshell, float r_cutoff, const double *r, int32_t n_r,
                           int32_t precision, double *energy);

#ifdef __cplusplus
}
#endif
#endif
