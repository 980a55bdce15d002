/*
 * CPU oracle for the CVF bulk-water Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (cvfbulkwater_amd/) never calls it.
 *
 * It restates, serially and in the reference's own storage layout (arm-major
 * uint8 sigma[6][N] and eta[6][N], a cube-state array, vertex/edge tables),
 * the update rules of lcoronas/CVFBulkWater SOURCE/gpu_water3D.cu with the
 * reference's cuRAND / MT19937 / xorshift streams replaced by one
 * counter-based generator (Philox4x32-10, Salmon et al., SC'11).  The
 * counter layout is the contract shared with the HIP engine:
 *
 *     counter = (site, step_lo32, (stream << 24) | sub, step_hi32)
 *     key     = (seed_lo32, seed_hi32)
 *
 * Pinning status (each function's comment names its checks):
 *   philox            KATs from the Random123 distribution + rocRAND header
 *   shell table       SOURCE/distances (exact) + brute-force pair enumeration
 *   vdw energy        closed form, pair sum == shell sum, data.out identities
 *   measure           data.out per-row identities, water3D.log summary lines
 *   cluster stats     SOURCE/cluster.out degenerate row, brute-force sizes
 *   sigma/eta/volume  exact single-variable transition probabilities vs.
 *                     empirical frequencies; long-run means vs SOURCE/data.out
 *   SW labels         BFS on small lattices (canonical minimum-node labels)
 *   non-chessboard    nhb bookkeeping invariant, exact single-variable
 *                     transition probabilities of the NHBMAX-gated rule, the
 *                     greedy capped N_HB count vs a brute-force restatement
 *                     on small lattices
 * The exact per-seed trajectories of the reference depend on cuRAND XORWOW
 * streams and cannot be reproduced: the oracle's trajectories are pinned to
 * the reference only statistically (parity unpinned bitwise at that
 * boundary, by design).
 */
#ifndef CVF_ORACLE_H
#define CVF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORA_ST_SCHED = 1, ORA_ST_SIGMA = 2, ORA_ST_ETA = 3, ORA_ST_VOLUME = 4,
  ORA_ST_SWBOND = 5, ORA_ST_SWFLIP = 6, ORA_ST_INIT_SIGMA = 7, ORA_ST_INIT_CUBE = 8, ORA_ST_SIGMA_F = 9
};

typedef struct ora_params {
  int32_t L;
  float P, T;
  uint64_t seed;
  int32_t n_metropolis, n_cluster;
  float r_cutoff;
  float J_hb, J_sig, v_hb;
  int32_t flag_config;     /* 0 random, 1 from arrays (ora_set_config), 2 ordered */
  int32_t flag_chessboard; /* 1: eta cube moves; 0: eta fixed, HB gated by NHBMAX */
  int32_t schedule;        /* 0: the reference's half-sweeps; 1: two-colour cell passes */
} ora_params;

typedef struct ora_obs {
  double E, H, V, n_hb, m, r_cell;
  int64_t nhb2;            /* 2 * N_HB, exact */
  int64_t n_sigma;         /* sum over cells of equal sigma pairs, exact */
} ora_obs;

typedef struct ora_cstat {
  int32_t largest, second, n_perc, n_clusters;
  float mean, var;
} ora_cstat;

typedef struct ora_sim ora_sim;

/* Philox4x32-10 block function. */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Shell table of gpu_water3D.cu:865-915 for a cubic L lattice.  Returns the
 * number of shells written (<= cap). */
int ora_shell_table(int L, float r_cutoff, float *dist, int32_t *freq, int cap);

/* E_vdW of gpu_water3D.cu:1784-1805 for a given shell table. */
double ora_vdw_energy(double r_cell, float r_cutoff, const float *dist,
                      const int32_t *freq, int nshell);

ora_sim *ora_create(const ora_params *p);
void ora_destroy(ora_sim *s);
/* sigma6N / eta6N arm-major (index arm*N + cell), values as in input_config. */
int ora_set_config(ora_sim *s, const uint8_t *sigma6N, const uint8_t *eta6N);
void ora_get_config(const ora_sim *s, uint8_t *sigma6N, uint8_t *eta6N);
void ora_get_cube_states(const ora_sim *s, int32_t *states);
/* per-cell HB counter nhb[N] of the non-chessboard mode */
void ora_get_nhb(const ora_sim *s, uint8_t *nhbN);

/* One full MC step of schedule R (gpu_water3D.cu:1346-1733). */
void ora_step(ora_sim *s);
void ora_steps(ora_sim *s, int64_t n);
/* Individual phases, for fine-grained parity tests. */
void ora_eta_sweep(ora_sim *s, uint64_t step);
void ora_sigma_halfsweep(ora_sim *s, uint64_t step, uint32_t sub, int arm, int indices);
/* colour pass of the fast schedule: pass = 2 (Metropolis iteration) + colour */
void ora_sigma_colour(ora_sim *s, uint64_t step, uint32_t pass);
void ora_schedule(uint64_t seed, uint64_t step, uint32_t sub, int *arm, int *indices);
void ora_volume_move(ora_sim *s, uint64_t step);
void ora_sw_step(ora_sim *s, uint64_t step, uint32_t jc);

uint64_t ora_get_step(const ora_sim *s);
void ora_set_step(ora_sim *s, uint64_t step);
void ora_measure(const ora_sim *s, ora_obs *o);
void ora_get_labels(const ora_sim *s, int32_t *labels6N);
int ora_cluster_stats(const ora_sim *s, ora_cstat *c);
void ora_get_volume_state(const ora_sim *s, double *r_cell, double *V, double *E_vdw);
void ora_get_thresholds(const ora_sim *s, uint64_t *thr_sigma33, uint64_t *thr_eta9,
                        uint64_t *thr_sw2);

/* Cluster statistics of an arbitrary label array (gpu_water3D.cu:2677-2780). */
int ora_cluster_stats_labels(int L, const int32_t *labels6N, ora_cstat *c);

/* Output lines in the reference formats. */
int ora_format_data_line(char *buf, int cap, int t, const ora_obs *o);
int ora_format_cluster_line(char *buf, int cap, int t, const ora_cstat *c);

#ifdef __cplusplus
}
#endif
#endif
