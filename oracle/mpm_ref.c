/*
 * Oracle MLS-MPM substep (see orc_mpm.h).  TEST INFRASTRUCTURE ONLY.
 * The substep sequence follows MPMSolver.substep (mpm_exps/engine/mpm_solver.py:624-631):
 * clear grid, P2G, grid update, G2P.  Build with -ffp-contract=off.
 */
#include "orc_mpm.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL float
#define SFX f32
#define FLOOR floorf
#define SQRT sqrtf
#include "orc_mpm_impl.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef SQRT

#define REAL double
#define SFX f64
#define FLOOR floor
#define SQRT sqrt
#include "orc_mpm_impl.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef SQRT

int orc_num_comps(const orc_mpm *s) {
  int d = s->dim;
  return 2 * d + d * d + (s->material == ORC_FLUID ? 1 : d * d);
}

static int64_t num_nodes(const orc_mpm *s) {
  int64_t n = s->n_grid;
  return s->dim == 3 ? n * n * n : n * n;
}

/* P2G split over threads: each thread scatters a contiguous particle range
   into a private grid, the grids are then summed node by node.  Used only to
   time the CPU baseline; the single-thread path is the parity oracle. */
#define DEFINE_PAR_P2G(REALT, SFXN)                                               \
  static void par_p2g_##SFXN(const orc_mpm *s, const REALT *st, int64_t n,        \
                             REALT *gm, REALT *gv, int nthreads) {                \
    int64_t nodes = num_nodes(s);                                                 \
    int d = s->dim;                                                               \
    if (nthreads <= 1) {                                                          \
      orc_p2g_##SFXN(s, st, n, gm, gv);                                           \
      return;                                                                     \
    }                                                                             \
    REALT *priv = (REALT *)calloc((size_t)nthreads * (size_t)nodes * (size_t)(d + 1), \
                                  sizeof(REALT));                                 \
    _Pragma("omp parallel num_threads(nthreads)") {                               \
      int t = 0, nt = 1;                                                          \
      SET_T                                                                       \
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt;                             \
      REALT *pm = priv + (size_t)t * (size_t)nodes * (size_t)(d + 1);             \
      REALT *pv = pm + nodes;                                                     \
      p2g_range_##SFXN(s, st, n, lo, hi, pm, pv);                                 \
    }                                                                             \
    _Pragma("omp parallel for num_threads(nthreads)")                             \
    for (int64_t i = 0; i < nodes; i++) {                                         \
      for (int c = 0; c <= d; c++) {                                              \
        REALT acc = 0;                                                            \
        for (int t = 0; t < nthreads; t++)                                        \
          acc += priv[(size_t)t * (size_t)nodes * (size_t)(d + 1) + (size_t)c * nodes + i]; \
        if (c == 0) gm[i] += acc; else gv[(int64_t)(c - 1) * nodes + i] += acc;   \
      }                                                                           \
    }                                                                             \
    free(priv);                                                                   \
  }

#ifdef _OPENMP
#define SET_T t = omp_get_thread_num(); nt = omp_get_num_threads();
#else
#define SET_T
#endif

/* P2G over particles [lo, hi): the oracle body with a sub-range.  Implemented
   by offsetting: the phase functions take (state, n) with stride n, so a
   range is handled by a view that keeps the stride. */
#define DEFINE_RANGE(REALT, SFXN)                                                 \
  static void p2g_range_##SFXN(const orc_mpm *s, const REALT *st, int64_t n,      \
                               int64_t lo, int64_t hi, REALT *gm, REALT *gv) {    \
    int nc = orc_num_comps(s);                                                    \
    int64_t m = hi - lo;                                                          \
    if (m <= 0) return;                                                           \
    REALT *tmp = (REALT *)malloc(sizeof(REALT) * (size_t)nc * (size_t)m);         \
    for (int c = 0; c < nc; c++)                                                  \
      memcpy(tmp + (size_t)c * m, st + (size_t)c * n + lo, sizeof(REALT) * (size_t)m); \
    orc_p2g_##SFXN(s, tmp, m, gm, gv);                                            \
    free(tmp);                                                                    \
  }

DEFINE_RANGE(float, f32)
DEFINE_RANGE(double, f64)
DEFINE_PAR_P2G(float, f32)
DEFINE_PAR_P2G(double, f64)

#define DEFINE_PAR_G2P(REALT, SFXN)                                               \
  static void par_g2p_##SFXN(const orc_mpm *s, const REALT *st, const REALT *gv,  \
                             REALT *out, int64_t n, int nthreads) {               \
    if (nthreads <= 1) {                                                          \
      orc_g2p_##SFXN(s, st, gv, out, n);                                          \
      return;                                                                     \
    }                                                                             \
    int nc = orc_num_comps(s);                                                    \
    _Pragma("omp parallel num_threads(nthreads)") {                               \
      int t = 0, nt = 1;                                                          \
      SET_T                                                                       \
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt, m = hi - lo;                \
      if (m > 0) {                                                                \
        REALT *a = (REALT *)malloc(sizeof(REALT) * (size_t)nc * (size_t)m * 2);   \
        REALT *b = a + (size_t)nc * m;                                            \
        for (int c = 0; c < nc; c++)                                              \
          memcpy(a + (size_t)c * m, st + (size_t)c * n + lo, sizeof(REALT) * (size_t)m); \
        orc_g2p_##SFXN(s, a, gv, b, m);                                           \
        for (int c = 0; c < nc; c++)                                              \
          memcpy(out + (size_t)c * n + lo, b + (size_t)c * m, sizeof(REALT) * (size_t)m); \
        free(a);                                                                  \
      }                                                                           \
    }                                                                             \
  }

DEFINE_PAR_G2P(float, f32)
DEFINE_PAR_G2P(double, f64)

void orc_run_plain_f32(const orc_mpm *s, float *state, int64_t n, int nsteps,
                       float *grid_m, float *grid_v, int nthreads) {
  int nc = orc_num_comps(s);
  float *out = (float *)malloc(sizeof(float) * (size_t)nc * (size_t)n);
  for (int t = 0; t < nsteps; t++) {
    zero_grid_f32(s, grid_m, grid_v);
    par_p2g_f32(s, state, n, grid_m, grid_v, nthreads);
    orc_grid_op_f32(s, grid_m, grid_v);
    par_g2p_f32(s, state, grid_v, out, n, nthreads);
    memcpy(state, out, sizeof(float) * (size_t)nc * (size_t)n);
  }
  free(out);
}

void orc_run_plain_f64(const orc_mpm *s, double *state, int64_t n, int nsteps,
                       double *grid_m, double *grid_v, int nthreads) {
  int nc = orc_num_comps(s);
  double *out = (double *)malloc(sizeof(double) * (size_t)nc * (size_t)n);
  for (int t = 0; t < nsteps; t++) {
    zero_grid_f64(s, grid_m, grid_v);
    par_p2g_f64(s, state, n, grid_m, grid_v, nthreads);
    orc_grid_op_f64(s, grid_m, grid_v);
    par_g2p_f64(s, state, grid_v, out, n, nthreads);
    memcpy(state, out, sizeof(double) * (size_t)nc * (size_t)n);
  }
  free(out);
}

void orc_run_quant_f32(const orc_mpm *s, const orc_comp *comps, int nwords,
                       uint64_t *words, int64_t n, int nsteps, uint64_t seed,
                       uint32_t step0, float *grid_m, float *grid_v,
                       float *scratch, int nthreads) {
  int nc = orc_num_comps(s);
  float *st = scratch, *out = scratch + (size_t)nc * (size_t)n;
  for (int t = 0; t < nsteps; t++) {
    int nt = nthreads < 1 ? 1 : nthreads;
    _Pragma("omp parallel for num_threads(nt)")
    for (int k = 0; k < nt; k++)
      orc_unpack_range_f32(comps, nc, nwords, words, st, n, n * k / nt, n * (k + 1) / nt);
    zero_grid_f32(s, grid_m, grid_v);
    par_p2g_f32(s, st, n, grid_m, grid_v, nthreads);
    orc_grid_op_f32(s, grid_m, grid_v);
    par_g2p_f32(s, st, grid_v, out, n, nthreads);
    _Pragma("omp parallel for num_threads(nt)")
    for (int k = 0; k < nt; k++)
      orc_pack_range_f32(comps, nc, nwords, out, NULL, seed, step0 + (uint32_t)t, words, n,
                         n * k / nt, n * (k + 1) / nt);
  }
}

void orc_limits_f32(const orc_mpm *s, const float *state, int64_t n, float *mx, float *mn) {
  int nc = orc_num_comps(s);
  for (int c = 0; c < nc; c++) {
    float a = -INFINITY, b = INFINITY;
    for (int64_t p = 0; p < n; p++) {
      float v = state[(int64_t)c * n + p];
      if (v > a) a = v;
      if (v < b) b = v;
    }
    mx[c] = a;
    mn[c] = b;
  }
}
