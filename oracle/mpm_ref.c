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
#define LOG logf
#define FMAX fmaxf
#define FMIN fminf
#include "orc_mpm_impl.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef SQRT
#undef LOG
#undef FMAX
#undef FMIN

#define REAL double
#define SFX f64
#define FLOOR floor
#define SQRT sqrt
#define LOG log
#define FMAX fmax
#define FMIN fmin
#include "orc_mpm_impl.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef SQRT
#undef LOG
#undef FMAX
#undef FMIN

int orc_num_comps(const orc_mpm *s) {
  int d = s->dim;
  return 2 * d + d * d + (s->material == ORC_FLUID ? 1 : d * d);
}

static int64_t num_nodes(const orc_mpm *s) {
  int64_t n = s->n_grid;
  return s->dim == 3 ? n * n * n : n * n;
}

#ifdef _OPENMP
#define THREAD_ID() omp_get_thread_num()
#define THREAD_COUNT() omp_get_num_threads()
#else
#define THREAD_ID() 0
#define THREAD_COUNT() 1
#endif

/* Multi-threaded phases, used only to time the CPU baseline (the
   single-thread path is the parity oracle).  P2G: each thread scatters a
   contiguous particle range into a private grid kept across calls; the
   private grids are then summed node by node. */
static void *g_priv = NULL;
static size_t g_priv_bytes = 0;

static void *priv_buffer(size_t bytes) {
  if (bytes > g_priv_bytes) {
    free(g_priv);
    g_priv = malloc(bytes);
    g_priv_bytes = g_priv ? bytes : 0;
  }
  return g_priv;
}

#define DEFINE_PAR(REALT, SFXN)                                                             \
  static void par_p2g_##SFXN(const orc_mpm *s, const REALT *st, int64_t n, REALT *gm,       \
                             REALT *gv, int nthreads) {                                     \
    const int64_t nodes = num_nodes(s);                                                     \
    const int d = s->dim;                                                                   \
    if (nthreads <= 1) {                                                                    \
      orc_p2g_##SFXN(s, st, n, gm, gv);                                                     \
      return;                                                                               \
    }                                                                                       \
    const size_t per = (size_t)nodes * (size_t)(d + 1);                                     \
    REALT *priv = (REALT *)priv_buffer((size_t)nthreads * per * sizeof(REALT));             \
    _Pragma("omp parallel num_threads(nthreads)") {                                         \
      const int t = THREAD_ID(), nt = THREAD_COUNT();                                       \
      REALT *pm = priv + (size_t)t * per;                                                   \
      memset(pm, 0, per * sizeof(REALT));                                                   \
      p2g_span_##SFXN(s, st, n, n * t / nt, n * (t + 1) / nt, pm, pm + nodes);              \
    }                                                                                       \
    _Pragma("omp parallel for num_threads(nthreads)")                                       \
    for (int64_t i = 0; i < nodes; i++) {                                                   \
      for (int c = 0; c <= d; c++) {                                                        \
        REALT acc = 0;                                                                      \
        for (int t = 0; t < nthreads; t++) acc += priv[(size_t)t * per + (size_t)c * nodes + i]; \
        if (c == 0) gm[i] += acc; else gv[(int64_t)(c - 1) * nodes + i] += acc;             \
      }                                                                                     \
    }                                                                                       \
  }                                                                                         \
  static void par_g2p_##SFXN(const orc_mpm *s, const REALT *st, const REALT *gv, REALT *out, \
                             int64_t n, int nthreads) {                                     \
    if (nthreads <= 1) {                                                                    \
      orc_g2p_##SFXN(s, st, gv, out, n);                                                    \
      return;                                                                               \
    }                                                                                       \
    _Pragma("omp parallel num_threads(nthreads)") {                                         \
      const int t = THREAD_ID(), nt = THREAD_COUNT();                                       \
      g2p_span_##SFXN(s, st, gv, out, n, n * t / nt, n * (t + 1) / nt);                     \
    }                                                                                       \
  }

DEFINE_PAR(float, f32)
DEFINE_PAR(double, f64)

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
  const int nt = nthreads < 1 ? 1 : nthreads;
  for (int t = 0; t < nsteps; t++) {
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

/* whole substeps on quantized storage with f64 compute (the reference's 2D
   quantized runs, init_quantaichi with compute_fp f64): decode in double,
   substep in double, encode in double (orc_pack_f64). */
void orc_run_quant_f64(const orc_mpm *s, const orc_comp *comps, int nwords,
                       uint64_t *words, int64_t n, int nsteps, uint64_t seed,
                       uint32_t step0, double *grid_m, double *grid_v,
                       double *scratch, int nthreads) {
  int nc = orc_num_comps(s);
  double *st = scratch, *out = scratch + (size_t)nc * (size_t)n;
  for (int t = 0; t < nsteps; t++) {
    orc_unpack_f64(comps, nc, nwords, words, st, n);
    zero_grid_f64(s, grid_m, grid_v);
    par_p2g_f64(s, st, n, grid_m, grid_v, nthreads);
    orc_grid_op_f64(s, grid_m, grid_v);
    par_g2p_f64(s, st, grid_v, out, n, nthreads);
    orc_pack_f64(comps, nc, nwords, out, NULL, seed, step0 + (uint32_t)t, words, n);
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
