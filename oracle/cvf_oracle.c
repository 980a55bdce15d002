/*
 * CPU oracle for the CVF bulk-water Monte Carlo hot path (test
 * infrastructure only; see cvf_oracle.h for the scope and pinning notes).
 *
 * Storage follows the reference, not the engine: sigma and eta are
 * arm-major uint8 arrays (index arm*N + cell, gpu_water3D.cu:530-536),
 * cubes keep an explicit state array and vertex table
 * (gpu_water3D.cu:3180-3257), and cluster labels are int32[6N] with node
 * id = arm*N + cell (gpu_water3D.cu:2376-2378).  Everything is serial.
 *
 * Build with -ffp-contract=off so that no multiply-add is fused.
 */
#include "cvf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ARMS 6
#define EDGES 12
#define NSOL 9
#define SHELL_CAP 256

struct ora_sim {
  ora_params p;
  int L, N, NC;
  uint8_t *s;        /* sigma [6][N] */
  uint8_t *eta;      /* eta   [6][N] */
  uint8_t *nhb;      /* nhb [N], non-chessboard mode (gpu_water3D.cu:125, 845-862) */
  int32_t *state;    /* cube state [N/4] */
  int32_t *vert;     /* [2][12][N/4] cube edge end cells */
  int32_t *label;    /* [6N] */
  int32_t *uf;       /* [6N] union-find scratch */
  double r_cell, V, E_vdw;
  int nshell;
  float dist[SHELL_CAP];
  int32_t freq[SHELL_CAP];
  double Jeff, beta;
  uint64_t thr_sigma[33], thr_eta[9], thr_hb, thr_s;
  uint64_t step;
};

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
 * as easy as 1, 2, 3").  Pinned by the Random123 known-answer vectors. */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void draw(uint64_t seed, uint32_t stream, uint32_t site, uint64_t step,
                 uint32_t sub, uint32_t out[4]) {
  uint32_t ctr[4] = {site, (uint32_t)step, (stream << 24) | sub, (uint32_t)(step >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  ora_philox(ctr, key, out);
}

/* floor(x * n / 2^32): an exact integer form of (int)(n*u), u in [0,1). */
static uint32_t uniform_int(uint32_t x, uint32_t n) {
  return (uint32_t)(((uint64_t)x * n) >> 32);
}

/* 53-bit uniform in [0,1) from two words. */
static double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* One 32-bit word per site update: with X = q*w (35-bit), the proposal is
 * X >> 32 = floor(q*u) and the acceptance uniform is the fractional part
 * X mod 2^32 (resolution q / 2^32), u = w / 2^32. */
static void split_word(uint32_t w, uint32_t q, uint32_t *prop, uint32_t *acc) {
  uint64_t X = (uint64_t)w * q;
  *prop = (uint32_t)(X >> 32);
  *acc = (uint32_t)X;
}

static uint64_t thr_from_p(double p) {
  if (!(p < 1.0)) return (uint64_t)1 << 32;
  if (p <= 0.0) return 0;
  return (uint64_t)floor(ldexp(p, 32));
}

/* ---------------------------------------------------------------- geometry */
static void coords(int L, int c, int *x, int *y, int *z) {
  *x = c % L;
  *y = (c / L) % L;
  *z = c / (L * L);
}

/* neighbour of cell c along arm a (0:+x 1:-x 2:+y 3:-y 4:+z 5:-z),
 * gpu_water3D.cu:1936-1962. */
static int neighbor(int L, int c, int a) {
  int x, y, z;
  coords(L, c, &x, &y, &z);
  switch (a) {
    case 0: x = (x + 1) % L; break;
    case 1: x = (x - 1 + L) % L; break;
    case 2: y = (y + 1) % L; break;
    case 3: y = (y - 1 + L) % L; break;
    case 4: z = (z + 1) % L; break;
    default: z = (z - 1 + L) % L; break;
  }
  return x + L * y + L * L * z;
}

/* Cube solution table, chessboard.h:211-262 (1 = active edge). */
static const int sol_off[NSOL][4] = {
    {2, 3, 6, 9}, {2, 3, 5, 11}, {0, 6, 8, 9}, {0, 5, 8, 11}, {1, 3, 7, 11},
    {1, 4, 7, 10}, {2, 4, 5, 10}, {1, 4, 8, 9}, {0, 6, 7, 10}};
static int Sol[NSOL][EDGES];

static void init_solutions(void) {
  for (int s = 0; s < NSOL; ++s) {
    for (int j = 0; j < EDGES; ++j) Sol[s][j] = 1;
    for (int k = 0; k < 4; ++k) Sol[s][sol_off[s][k]] = 0;
  }
}

/* arm pair of edge j: x-edges {0,6,8,9}, y-edges {1,4,7,10}, z-edges
 * {2,3,5,11} (gpu_water3D.cu:3422-3431). */
static int edge_arm0(int j) {
  if (j == 0 || j == 6 || j == 8 || j == 9) return 0;
  if (j == 1 || j == 4 || j == 7 || j == 10) return 2;
  return 4;
}

/* Cube vertex table, gpu_water3D.cu:3180-3257. */
static void build_cubes(ora_sim *s) {
  int L = s->L, NC = s->NC;
  for (int c = 0; c < s->N; ++c) {
    int x, y, z;
    coords(L, c, &x, &y, &z);
    if (z % 2 == 0) {
      if (x % 2 != 0 || y % 2 != 0) continue;
    } else {
      if (x % 2 != 1 || y % 2 != 1) continue;
    }
    int cube = x / 2 + (L / 2) * (y / 2) + (L * L / 4) * z;
    int xp = (x + 1) % L, yp = (y + 1) % L, zp = (z + 1) % L;
    int v1 = x + y * L + z * L * L;
    int v2 = xp + y * L + z * L * L;
    int v3 = x + yp * L + z * L * L;
    int v4 = x + y * L + zp * L * L;
    int v6 = x + yp * L + zp * L * L;
    int v7 = xp + y * L + zp * L * L;
    int v8 = xp + yp * L + z * L * L;
    int v5 = xp + yp * L + zp * L * L;
    int a[EDGES] = {v1, v1, v1, v2, v2, v3, v3, v4, v4, v6, v7, v8};
    int b[EDGES] = {v2, v3, v4, v7, v8, v6, v8, v6, v7, v5, v5, v5};
    for (int j = 0; j < EDGES; ++j) {
      s->vert[cube + j * NC] = a[j];
      s->vert[cube + j * NC + NC * EDGES] = b[j];
    }
  }
}

static void write_cube_edges(ora_sim *s, int cube) {
  int N = s->N, NC = s->NC, st = s->state[cube];
  for (int j = 0; j < EDGES; ++j) {
    int v0 = s->vert[cube + j * NC], v1 = s->vert[cube + j * NC + NC * EDGES];
    int a0 = edge_arm0(j);
    s->eta[a0 * N + v0] = (uint8_t)Sol[st][j];
    s->eta[(a0 + 1) * N + v1] = (uint8_t)Sol[st][j];
  }
}

/* Decision tree of gpu_Chessboard_set_state, gpu_water3D.cu:3314-3371. */
static int cube_state_from_eta(const ora_sim *s, int cube) {
  int N = s->N, NC = s->NC;
  const int32_t *V = s->vert;
  int v1 = V[cube + 0 * NC], v2 = V[cube + 3 * NC], v3 = V[cube + 5 * NC];
  int v6 = V[cube + 5 * NC + NC * EDGES], v7 = V[cube + 3 * NC + NC * EDGES];
  int v8 = V[cube + 4 * NC + NC * EDGES];
  const uint8_t *e = s->eta;
  if (e[0 * N + v1] == 0) {
    if (e[0 * N + v6] == 0) return 2;
    if (e[4 * N + v3] == 0) return 3;
    return 8;
  } else if (e[2 * N + v1] == 0) {
    if (e[4 * N + v2] == 0) return 4;
    if (e[2 * N + v7] == 0) return 5;
    return 7;
  }
  if (e[0 * N + v3] == 0) return 0;
  if (e[4 * N + v8] == 0) return 1;
  return 6;
}

/* ------------------------------------------------------------ shell table */
/* distance_PBC, gpu_water3D.cu:1323-1337 (float result of a double sqrt). */
static float distance_pbc(int L, int x1, int y1, int z1, int x2, int y2, int z2) {
  int xp = abs(x1 - x2), yp = abs(y1 - y2), zp = abs(z1 - z2);
  if (xp > L / 2) xp = L - xp;
  if (yp > L / 2) yp = L - yp;
  if (zp > L / 2) zp = L - zp;
  return (float)sqrt((double)(xp * xp + yp * yp + zp * zp));
}

/* gpu_water3D.cu:865-915: distinct float distances from cell 0 below the
 * cutoff in discovery order, multiplicity halved with integer division and
 * multiplied by N. */
int ora_shell_table(int L, float r_cutoff, float *dist, int32_t *freq, int cap) {
  int N = L * L * L, n = 0;
  for (int j = 1; j < N; ++j) {
    int x2 = j % L, y2 = (j / L) % L, z2 = j / (L * L);
    float d = distance_pbc(L, 0, 0, 0, x2, y2, z2);
    if (d < r_cutoff) {
      int k;
      for (k = 0; k < n; ++k)
        if (d == dist[k]) {
          freq[k]++;
          break;
        }
      if (k == n) {
        if (n == cap) return -1;
        dist[n] = d;
        freq[n] = 1;
        ++n;
      }
    }
  }
  for (int k = 0; k < n; ++k) {
    freq[k] /= 2;
    freq[k] *= N;
  }
  return n;
}

/* calculate_energy_vdW, gpu_water3D.cu:1784-1805. */
double ora_vdw_energy(double r_cell, float r_cutoff, const float *dist,
                      const int32_t *freq, int nshell) {
  double tmp = 1.0 / (r_cutoff);
  tmp = tmp * tmp * tmp;
  double attractive = tmp * tmp;
  double repulsive = attractive * attractive;
  double shift = repulsive - attractive;
  double ret = 0;
  for (int i = 0; i < nshell; ++i)
    if (r_cell * dist[i] < r_cutoff) {
      tmp = 1.0 / (r_cell * dist[i]);
      tmp = tmp * tmp * tmp;
      attractive = tmp * tmp;
      repulsive = attractive * attractive;
      ret += freq[i] * (repulsive - attractive - shift);
    }
  return ret;
}

/* ------------------------------------------------------------ life cycle */
static void compute_thresholds(ora_sim *s) {
  const ora_params *p = &s->p;
  /* Jeff = J_hb - P*v_hb in float arithmetic (gpu_water3D.cu:391, 76). */
  float jf = p->J_hb - p->P * p->v_hb;
  s->Jeff = (double)jf;
  s->beta = 1. / p->T; /* gpu_water3D.cu:388 */
  for (int h = -1; h <= 1; ++h)
    for (int d = -5; d <= 5; ++d) {
      double dE = -s->Jeff * h - (double)p->J_sig * d;
      s->thr_sigma[(h + 1) * 11 + (d + 5)] = thr_from_p(exp(-(s->beta * dE)));
    }
  for (int dn = -4; dn <= 4; ++dn)
    s->thr_eta[dn + 4] = thr_from_p(exp(s->beta * (s->Jeff * (double)dn)));
  /* gpu_water3D.cu:992-993 */
  s->thr_hb = thr_from_p(1.0 - exp(-s->beta * fabs(s->Jeff)));
  s->thr_s = thr_from_p(1.0 - exp(-s->beta * fabs((double)p->J_sig)));
}

/* Initial HB counters, gpu_water3D.cu:845-862: per cell, the number of
 * arms whose sigma equals the facing arm of the neighbour, weighted by eta
 * (uint8 arithmetic). */
static void compute_nhb(ora_sim *s) {
  int L = s->L, N = s->N;
  for (int c = 0; c < N; ++c) {
    uint8_t acc = 0;
    for (int a = 0; a < ARMS; ++a) {
      int nb = neighbor(L, c, a);
      acc = (uint8_t)(acc + s->eta[a * N + c] * (s->s[a * N + c] == s->s[(a ^ 1) * N + nb]));
    }
    s->nhb[c] = acc;
  }
}

ora_sim *ora_create(const ora_params *p) {
  if (p->L < 4 || p->L % 4 != 0) return NULL;
  /* the colour passes need the chessboard rule: a counter update touches a
   * neighbour that other cells of the pass also touch */
  if (p->schedule != 0 && (p->schedule != 1 || !p->flag_chessboard)) return NULL;
  init_solutions();
  ora_sim *s = (ora_sim *)calloc(1, sizeof(ora_sim));
  s->p = *p;
  s->L = p->L;
  s->N = p->L * p->L * p->L;
  s->NC = s->N / 4;
  s->s = (uint8_t *)calloc((size_t)ARMS * s->N, 1);
  s->eta = (uint8_t *)calloc((size_t)ARMS * s->N, 1);
  s->nhb = (uint8_t *)calloc((size_t)s->N, 1);
  s->state = (int32_t *)calloc((size_t)s->NC, sizeof(int32_t));
  s->vert = (int32_t *)calloc((size_t)2 * EDGES * s->NC, sizeof(int32_t));
  s->label = (int32_t *)calloc((size_t)ARMS * s->N, sizeof(int32_t));
  s->uf = (int32_t *)calloc((size_t)ARMS * s->N, sizeof(int32_t));
  compute_thresholds(s);
  build_cubes(s);
  s->nshell = ora_shell_table(s->L, p->r_cutoff, s->dist, s->freq, SHELL_CAP);
  int N = s->N;
  /* initial sigma (gpu_water3D.cu:812-826) with Philox replacing MT19937 */
  if (p->flag_config == 2) {
    uint32_t r[4];
    draw(p->seed, ORA_ST_INIT_SIGMA, 0, 0, 2, r);
    uint8_t v = (uint8_t)uniform_int(r[0], 6);
    memset(s->s, v, (size_t)ARMS * N);
  } else {
    for (int c = 0; c < N; ++c) {
      uint32_t r0[4], r1[4];
      draw(p->seed, ORA_ST_INIT_SIGMA, (uint32_t)c, 0, 0, r0);
      draw(p->seed, ORA_ST_INIT_SIGMA, (uint32_t)c, 0, 1, r1);
      uint32_t w[6] = {r0[0], r0[1], r0[2], r0[3], r1[0], r1[1]};
      for (int a = 0; a < ARMS; ++a) s->s[a * N + c] = (uint8_t)uniform_int(w[a], 6);
    }
  }
  if (p->flag_chessboard) {
    /* random initial cube states (gpu_Chessboard_start_cubes, 3259-3312) */
    for (int cube = 0; cube < s->NC; ++cube) {
      uint32_t r[4];
      draw(p->seed, ORA_ST_INIT_CUBE, (uint32_t)cube, 0, 0, r);
      s->state[cube] = (int32_t)uniform_int(r[0], 9);
      write_cube_edges(s, cube);
    }
  } else {
    /* every bond active (gpu_water3D.cu:721-724) */
    memset(s->eta, 1, (size_t)ARMS * N);
  }
  compute_nhb(s);
  /* volume chain (gpu_water3D.cu:150, 481, 1128) */
  s->r_cell = 1.0001;
  s->V = N * s->r_cell * s->r_cell * s->r_cell;
  s->E_vdw = ora_vdw_energy(s->r_cell, p->r_cutoff, s->dist, s->freq, s->nshell);
  s->step = 0;
  return s;
}

void ora_destroy(ora_sim *s) {
  if (!s) return;
  free(s->s); free(s->eta); free(s->nhb); free(s->state); free(s->vert); free(s->label); free(s->uf);
  free(s);
}

/* input_config semantics (gpu_water3D.cu:730-800): sigma and eta from the
 * caller, cube states from the eta decision tree; the cube edges are then
 * rewritten from the solution table, which is what the reference's first
 * eta sweep does for every cube. Returns the number of cubes whose eta did
 * not already match their solution, or -1 on an out-of-range value. */
int ora_set_config(ora_sim *s, const uint8_t *sigma6N, const uint8_t *eta6N) {
  int N = s->N;
  for (int i = 0; i < ARMS * N; ++i)
    if (sigma6N[i] > 5 || eta6N[i] > 1) return -1;
  memcpy(s->s, sigma6N, (size_t)ARMS * N);
  memcpy(s->eta, eta6N, (size_t)ARMS * N);
  if (!s->p.flag_chessboard) {
    /* eta is kept as read (no cube decoding, gpu_water3D.cu:794) */
    compute_nhb(s);
    return 0;
  }
  int bad = 0;
  for (int cube = 0; cube < s->NC; ++cube) {
    s->state[cube] = cube_state_from_eta(s, cube);
    int st = s->state[cube];
    for (int j = 0; j < EDGES; ++j) {
      int v0 = s->vert[cube + j * s->NC], v1 = s->vert[cube + j * s->NC + s->NC * EDGES];
      int a0 = edge_arm0(j);
      if (s->eta[a0 * N + v0] != Sol[st][j] || s->eta[(a0 + 1) * N + v1] != Sol[st][j]) {
        ++bad;
        break;
      }
    }
    write_cube_edges(s, cube);
  }
  compute_nhb(s);
  return bad;
}

void ora_get_config(const ora_sim *s, uint8_t *sigma6N, uint8_t *eta6N) {
  if (sigma6N) memcpy(sigma6N, s->s, (size_t)ARMS * s->N);
  if (eta6N) memcpy(eta6N, s->eta, (size_t)ARMS * s->N);
}

void ora_get_cube_states(const ora_sim *s, int32_t *states) {
  memcpy(states, s->state, sizeof(int32_t) * (size_t)s->NC);
}

void ora_get_nhb(const ora_sim *s, uint8_t *nhbN) { memcpy(nhbN, s->nhb, (size_t)s->N); }

uint64_t ora_get_step(const ora_sim *s) { return s->step; }
void ora_set_step(ora_sim *s, uint64_t step) { s->step = step; }

void ora_get_volume_state(const ora_sim *s, double *r_cell, double *V, double *E_vdw) {
  if (r_cell) *r_cell = s->r_cell;
  if (V) *V = s->V;
  if (E_vdw) *E_vdw = s->E_vdw;
}

void ora_get_thresholds(const ora_sim *s, uint64_t *thr_sigma33, uint64_t *thr_eta9,
                        uint64_t *thr_sw2) {
  if (thr_sigma33) memcpy(thr_sigma33, s->thr_sigma, sizeof(s->thr_sigma));
  if (thr_eta9) memcpy(thr_eta9, s->thr_eta, sizeof(s->thr_eta));
  if (thr_sw2) {
    thr_sw2[0] = s->thr_hb;
    thr_sw2[1] = s->thr_s;
  }
}

/* ------------------------------------------------------------ MC phases */
/* Cube Metropolis, gpu_Chessboard_set_active_bonds (gpu_water3D.cu:3374-3466):
 * proposal floor(9u), Delta N_HB over the 12 edges, accept with
 * min(1, exp(beta*Jeff*DeltaN)). */
void ora_eta_sweep(ora_sim *s, uint64_t step) {
  int N = s->N, NC = s->NC;
  for (int cube = 0; cube < NC; ++cube) {
    int old = s->state[cube];
    /* the four cubes with origins (ox + b L/2, oy + a L/2, oz), a,b in
     * {0,1}, share one Philox block keyed on the a=b=0 cube; cube a,b uses
     * word a + 2b (cube index = x/2 + (L/2)(y/2) + (L^2/4) z, 3405) */
    int L = s->L, hL = L / 2;
    int ox2 = cube % hL, oy2 = (cube / hL) % hL;
    int ay = oy2 >= L / 4, bx = ox2 >= L / 4;
    int site = cube - ay * hL * (L / 4) - bx * (L / 4);
    uint32_t r[4], prop, acc;
    draw(s->p.seed, ORA_ST_ETA, (uint32_t)site, step, 0, r);
    split_word(r[ay + 2 * bx], 9, &prop, &acc);
    int nw = (int)prop;
    int dn = 0;
    for (int j = 0; j < EDGES; ++j) {
      int v0 = s->vert[cube + j * NC], v1 = s->vert[cube + j * NC + NC * EDGES];
      int a0 = edge_arm0(j);
      dn += (Sol[nw][j] - Sol[old][j]) * (s->s[a0 * N + v0] == s->s[(a0 + 1) * N + v1]);
    }
    if (!((uint64_t)acc < s->thr_eta[dn + 4])) nw = old;
    s->state[cube] = nw;
    write_cube_edges(s, cube);
  }
}

/* (arm, parity) of a half-sweep: gpu_water3D.cu:1393-1394 draw
 * spin = floor(6u), indices = floor(2u) from MT19937; here from Philox. */
void ora_schedule(uint64_t seed, uint64_t step, uint32_t sub, int *arm, int *indices) {
  uint32_t r[4];
  draw(seed, ORA_ST_SCHED, 0, step, sub, r);
  *arm = (int)uniform_int(r[0], 6);
  *indices = (int)(r[1] >> 31);
}

/* gpu_update (gpu_water3D.cu:1890-2008): one arm of the cells whose
 * coordinate along the arm's axis is even (indices=1) or odd (indices=0).
 * Chessboard branch: the HB term is eta * (delta(new,nb) - delta(old,nb)).
 * Non-chessboard branch (1981-1984, 2004-2007): with dN = delta(new,nb) -
 * delta(old,nb), the HB term is [nhb_i+dN <= 4][nhb_j+dN <= 4] delta(new,nb)
 * - [nhb_i <= 4][nhb_j <= 4] delta(old,nb), and an accepted move adds dN to
 * the uint8 counters of both cells. */
/* One sigma move of arm `arm` of cell c to proposal `nw`, accepted iff
 * acc < threshold (the rule of gpu_update, 1966-2007, both branches). */
static void sigma_move(ora_sim *s, int c, int arm, int nw, uint32_t acc) {
  int L = s->L, N = s->N;
  int nb = neighbor(L, c, arm);
  int nbspin = s->s[(arm ^ 1) * N + nb];
  int old = s->s[arm * N + c];
  if (nw == old) return;
  int dn = (nw == nbspin) - (old == nbspin);
  int h;
  if (s->p.flag_chessboard) {
    h = s->eta[arm * N + c] * dn;
  } else {
    int ni = s->nhb[c], nj = s->nhb[nb];
    h = ((ni + dn <= 4) && (nj + dn <= 4)) * (nw == nbspin) - ((ni <= 4) && (nj <= 4)) * (old == nbspin);
  }
  int d = 0;
  for (int k = 0; k < ARMS; ++k) {
    if (k == arm) continue;
    int sk = s->s[k * N + c];
    d += (nw == sk) - (old == sk);
  }
  if ((uint64_t)acc < s->thr_sigma[(h + 1) * 11 + (d + 5)]) {
    s->s[arm * N + c] = (uint8_t)nw;
    s->nhb[c] = (uint8_t)(s->nhb[c] + dn);
    s->nhb[nb] = (uint8_t)(s->nhb[nb] + dn);
  }
}

/* the four cells (x + b L/2, y + a L/2, z), a,b in {0,1}, share parities
 * along every axis and the colour (L/2 is even) and one Philox block keyed
 * on the a=b=0 cell; cell a,b uses word a + 2b */
static int quad_site(int L, int c, int x, int y, int *word) {
  int ay = y >= L / 2, bx = x >= L / 2;
  *word = ay + 2 * bx;
  return c - ay * (L / 2) * L - bx * (L / 2);
}

void ora_sigma_halfsweep(ora_sim *s, uint64_t step, uint32_t sub, int arm, int indices) {
  int L = s->L, N = s->N;
  for (int c = 0; c < N; ++c) {
    int x, y, z;
    coords(L, c, &x, &y, &z);
    int coord = arm < 2 ? x : (arm < 4 ? y : z);
    if (indices && (coord % 2)) continue;
    if (!indices && !(coord % 2)) continue;
    int word, site = quad_site(L, c, x, y, &word);
    uint32_t r[4], prop, acc;
    draw(s->p.seed, ORA_ST_SIGMA, (uint32_t)site, step, sub, r);
    split_word(r[word], 6, &prop, &acc);
    sigma_move(s, c, arm, (int)prop, acc);
  }
}

/* The fast schedule's sigma pass (a design of this engine, not the
 * reference's schedule): every cell of colour (x + y + z) % 2 updates its
 * six arms in order 0..5 with the same single-move rule.  Cells of one
 * colour share no bond (every neighbour has the other colour), so the pass
 * is one parallel sweep.  Randoms: stream ORA_ST_SIGMA_F, the quad's site,
 * sub = 8 pass + arm, pass = 2 (Metropolis iteration) + colour. */
void ora_sigma_colour(ora_sim *s, uint64_t step, uint32_t pass) {
  int L = s->L, N = s->N, colour = (int)(pass & 1);
  for (int c = 0; c < N; ++c) {
    int x, y, z;
    coords(L, c, &x, &y, &z);
    if (((x + y + z) & 1) != colour) continue;
    int word, site = quad_site(L, c, x, y, &word);
    for (int arm = 0; arm < ARMS; ++arm) {
      uint32_t r[4], prop, acc;
      draw(s->p.seed, ORA_ST_SIGMA_F, (uint32_t)site, step, 8 * pass + (uint32_t)arm, r);
      split_word(r[word], 6, &prop, &acc);
      sigma_move(s, c, arm, (int)prop, acc);
    }
  }
}

/* volume_step, gpu_water3D.cu:1740-1781 (FP64, libm exp/log). */
void ora_volume_move(ora_sim *s, uint64_t step) {
  const ora_params *p = &s->p;
  int L = s->L, N = s->N;
  double trial, u2;
  uint32_t attempt = 0, r[4];
  double eps = 0.01;
  do {
    draw(p->seed, ORA_ST_VOLUME, attempt, step, 0, r);
    trial = s->r_cell + eps * 2 * (u53(r[0], r[1]) - 0.5);
    ++attempt;
  } while (trial < 0.0);
  u2 = u53(r[2], r[3]);
  double tv = trial * trial * trial * L * L * L;
  double dV = tv - s->V;
  double te = ora_vdw_energy(trial, p->r_cutoff, s->dist, s->freq, s->nshell);
  float NT = (float)N * p->T;
  double dE = (te - s->E_vdw) - NT * log(tv / s->V);
  if (u2 < exp(-s->beta * (dE + p->P * dV))) {
    s->r_cell = trial;
    s->V = tv;
    s->E_vdw = te;
  }
}

static int uf_find(int32_t *uf, int x) {
  int r = x;
  while (uf[r] != r) r = uf[r];
  while (uf[x] != r) {
    int n = uf[x];
    uf[x] = r;
    x = n;
  }
  return r;
}

static void uf_union(int32_t *uf, int a, int b) {
  a = uf_find(uf, a);
  b = uf_find(uf, b);
  if (a == b) return;
  if (a < b) uf[b] = a;
  else uf[a] = b;
}

/* Swendsen-Wang step: bond activation of gpu_create_cluster
 * (gpu_water3D.cu:2383-2438), canonical minimum-node labels (the fixed point
 * of 2440-2529), cluster rotation of gpu_update_cluster (2531-2542). */
void ora_sw_step(ora_sim *s, uint64_t step, uint32_t jc) {
  int L = s->L, N = s->N, NN = ARMS * N;
  uint8_t *S = s->s;
  for (int i = 0; i < NN; ++i) s->uf[i] = i;
  for (int c = 0; c < N; ++c) {
    uint32_t w[20];
    for (int g = 0; g < 5; ++g) draw(s->p.seed, ORA_ST_SWBOND, (uint32_t)c, step, jc * 5 + g, &w[4 * g]);
    int nbr[3] = {neighbor(L, c, 0), neighbor(L, c, 2), neighbor(L, c, 4)};
    for (int ax = 0; ax < 3; ++ax) {
      int a0 = 2 * ax, a1 = 2 * ax + 1, nb = nbr[ax];
      int act = 0;
      if (s->Jeff > 0)
        act = s->eta[a0 * N + c] && S[a0 * N + c] == S[a1 * N + nb] && (uint64_t)w[ax] < s->thr_hb;
      else if (s->Jeff < 0)
        act = S[a0 * N + c] != S[a1 * N + nb] && (uint64_t)w[ax] < s->thr_hb;
      if (act) uf_union(s->uf, a0 * N + c, a1 * N + nb);
    }
    int link = 3;
    for (int ii = 0; ii < ARMS - 1; ++ii)
      for (int jj = ii + 1; jj < ARMS; ++jj) {
        if (S[ii * N + c] == S[jj * N + c] && (uint64_t)w[link] < s->thr_s)
          uf_union(s->uf, ii * N + c, jj * N + c);
        ++link;
      }
  }
  for (int i = 0; i < NN; ++i) s->label[i] = uf_find(s->uf, i);
  for (int i = 0; i < NN; ++i) {
    uint32_t r[4];
    draw(s->p.seed, ORA_ST_SWFLIP, (uint32_t)s->label[i], step, jc, r);
    S[i] = (uint8_t)((S[i] + uniform_int(r[0], 6)) % 6);
  }
}

void ora_step(ora_sim *s) {
  uint64_t t = s->step;
  if (s->p.flag_chessboard) ora_eta_sweep(s, t);
  ora_volume_move(s, t);
  for (int j = 0; j < s->p.n_metropolis; ++j) {
    if (s->p.schedule == 1) {
      ora_sigma_colour(s, t, (uint32_t)(2 * j));
      ora_sigma_colour(s, t, (uint32_t)(2 * j + 1));
      continue;
    }
    for (int k = 0; k < 2 * ARMS; ++k) {
      int arm, ind;
      uint32_t sub = (uint32_t)(j * 2 * ARMS + k);
      ora_schedule(s->p.seed, t, sub, &arm, &ind);
      ora_sigma_halfsweep(s, t, sub, arm, ind);
    }
  }
  for (int jc = 0; jc < s->p.n_cluster; ++jc) ora_sw_step(s, t, (uint32_t)jc);
  s->step = t + 1;
}

void ora_steps(ora_sim *s, int64_t n) {
  for (int64_t i = 0; i < n; ++i) ora_step(s);
}

/* ------------------------------------------------------------ observables */
/* measure(), gpu_water3D.cu:2014-2160, same FP64 accumulation order.
 * Chessboard branch: every HB counts with its two eta.  Non-chessboard
 * branch (2064-2099): cells in index order, forward x, y, z bonds; an HB
 * counts for both cells only while both counters are below NHBMAX = 4. */
void ora_measure(const ora_sim *s, ora_obs *o) {
  int L = s->L, N = s->N;
  const uint8_t *S = s->s, *A = s->eta;
  int32_t *cnt = (int32_t *)calloc((size_t)N, sizeof(int32_t));
  double m = 0;
  int64_t nsig = 0;
  for (int i = 0; i < N; ++i) {
    for (int ax = 0; ax < 3; ++ax) {
      int a0 = 2 * ax, a1 = a0 + 1;
      int nb = neighbor(L, i, a0);
      if (S[a0 * N + i] == S[a1 * N + nb]) {
        if (s->p.flag_chessboard) {
          cnt[i] += A[a0 * N + i];
          cnt[nb] += A[a1 * N + nb];
        } else if (cnt[i] < 4 && cnt[nb] < 4) {
          ++cnt[i];
          ++cnt[nb];
        }
      }
    }
    for (int j = 0; j < ARMS - 1; ++j)
      for (int k = j + 1; k < ARMS; ++k)
        if (S[j * N + i] == S[k * N + i]) {
          m += 1;
          ++nsig;
        }
  }
  double e = ora_vdw_energy(s->r_cell, s->p.r_cutoff, s->dist, s->freq, s->nshell);
  double v = N * s->r_cell * s->r_cell * s->r_cell;
  double n_hb = 0;
  int64_t nhb2 = 0;
  for (int i = 0; i < N; ++i) {
    v += 0.5 * cnt[i] * s->p.v_hb;
    e += 0.5 * cnt[i] * (-s->p.J_hb);
    n_hb += 0.5 * cnt[i];
    nhb2 += cnt[i];
  }
  e += (-s->p.J_sig) * m;
  double h = e + s->p.P * v;
  m /= 1. * N * (ARMS * (ARMS - 1) / 2.);
  free(cnt);
  o->E = e;
  o->H = h;
  o->V = v;
  o->n_hb = n_hb;
  o->m = m;
  o->r_cell = s->r_cell;
  o->nhb2 = nhb2;
  o->n_sigma = nsig;
}

void ora_get_labels(const ora_sim *s, int32_t *labels6N) {
  memcpy(labels6N, s->label, sizeof(int32_t) * (size_t)ARMS * s->N);
}

/* print_cluster_statistics + is_percolating_cluster, gpu_water3D.cu:2677-2780:
 * float accumulation in ascending label order, int32 size*size that wraps. */
int ora_cluster_stats_labels(int L, const int32_t *label, ora_cstat *c) {
  int N = L * L * L, NN = ARMS * N;
  int32_t *size = (int32_t *)calloc((size_t)NN, sizeof(int32_t));
  uint8_t *X = (uint8_t *)calloc((size_t)L, 1), *Y = (uint8_t *)calloc((size_t)L, 1),
          *Z = (uint8_t *)calloc((size_t)L, 1);
  for (int i = 0; i < NN; ++i) {
    if (label[i] < 0 || label[i] >= NN) {
      free(size); free(X); free(Y); free(Z);
      return -1;
    }
    size[label[i]]++;
  }
  int largest = 0, second = 0, ncl = 0, nperc = 0;
  float average = 0, average2 = 0;
  for (int i = 0; i < NN; ++i) {
    if (size[i] == 0) continue;
    ncl++;
    average += size[i];
    average2 += (int32_t)((uint32_t)size[i] * (uint32_t)size[i]);
    if (size[i] > largest) {
      second = largest;
      largest = size[i];
    } else if (size[i] > second) {
      second = size[i];
    }
    if (size[i] >= 2 * L) {
      memset(X, 0, (size_t)L); memset(Y, 0, (size_t)L); memset(Z, 0, (size_t)L);
      int counter = 0;
      for (int k = 0; k < NN; ++k) {
        if (label[k] == i) {
          int cell = k % N;
          X[cell % L] = 1;
          Y[(cell / L) % L] = 1;
          Z[cell / (L * L)] = 1;
          counter++;
        }
        if (counter == size[i]) break;
      }
      int fx = 1, fy = 1, fz = 1;
      for (int k = 0; k < L; ++k) {
        if (!X[k]) fx = 0;
        if (!Y[k]) fy = 0;
        if (!Z[k]) fz = 0;
      }
      nperc += (fx || fy || fz);
    }
  }
  average /= (float)ncl;
  average2 /= (float)ncl;
  average2 -= average * average;
  c->largest = largest;
  c->second = second;
  c->n_perc = nperc;
  c->n_clusters = ncl;
  c->mean = average;
  c->var = average2;
  free(size); free(X); free(Y); free(Z);
  return 0;
}

int ora_cluster_stats(const ora_sim *s, ora_cstat *c) {
  return ora_cluster_stats_labels(s->L, s->label, c);
}

/* data.out row, gpu_water3D.cu:2159 */
int ora_format_data_line(char *buf, int cap, int t, const ora_obs *o) {
  return snprintf(buf, (size_t)cap, "%d %1.7e %1.7e %1.7e %1.7e %1.7e %1.7e\n", t, o->E, o->H,
                  o->V, o->n_hb, o->m, o->r_cell);
}

/* cluster.out row, gpu_water3D.cu:2717 */
int ora_format_cluster_line(char *buf, int cap, int t, const ora_cstat *c) {
  return snprintf(buf, (size_t)cap, "%d %d %d %d %f %f %d\n", t, c->largest, c->second,
                  c->n_perc, c->mean, c->var, c->n_clusters);
}
