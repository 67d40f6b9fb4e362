/*
 * CPU oracle for the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates MPMSolver.p2g_bls / grid_op / g2p_bls / compute_cauchy /
 * compute_ek / update_limits_* of the reference
 * (mpm_exps/engine/mpm_solver.py:578-758, 960-1052) and
 * polar_decompose2d (TS/python/taichi/lang/linalg_impl.py:9-26).
 *
 * State layout: component-major SoA, state[c * n + p], components in the
 * reference's bit_pack_assign order (mpm_solver.py:700-744):
 *   x[dim], v[dim], C[dim*dim] (row major), then J (fluid) or F[dim*dim].
 * Grid: dense n_grid^dim nodes, grid_m[node], grid_v[d * nodes + node],
 * node = (i * n + j) * n + k.
 */
#ifndef AQ_ORACLE_MPM_H
#define AQ_ORACLE_MPM_H
#include <stdint.h>
#include "orc.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FLUID = 0, ORC_ELASTIC = 1 };
enum { ORC_SEPARATING = 0, ORC_STICKY = 1 };

typedef struct orc_mpm {
  int32_t dim;       /* 2 or 3 */
  int32_t material;  /* ORC_FLUID (J, sigma = la (J-1) I) or ORC_ELASTIC
                        (2D fixed-corotated, sigma = 2 mu (F-R) F^T + la J (J-1) I;
                        3D sigma = mu (F F^T - I) + la log(J) on every entry) */
  int32_t n_grid;
  int32_t boundary;  /* ORC_SEPARATING (mpm_solver.py:659-664) or ORC_STICKY */
  int32_t bound;     /* wall thickness in cells: lb = bound, ub = n_grid - bound */
  int32_t pad;
  double dx, inv_dx, dt, p_vol, p_mass, gravity;
  double la;         /* fluid bulk modulus (MPMSolver.la) */
  double mu_0, la_0; /* elastic Lame parameters */
  /* MPMConfig.lb / ub (mpm_solver.py:97-105): axis_bounds = 1 puts the walls
     of axis a at lb[a], ub[a] instead of bound, n_grid - bound */
  int32_t axis_bounds, lb[3], ub[3];
  int32_t bound_clip; /* use_bound_clip (mpm_solver.py:696-698): G2P clamps x to [lb dx, ub dx] */
} orc_mpm;

/* The SDF collider at one substep (mpm_solver.py:1444-1545): the volume
   (entry (i, j, k) at node offset + (i, j, k)), the rotation centre, the
   rotation (cos, sin of theta), the height shift int(height / dx + 0.5) in
   nodes, the angular velocity, and whether the collider velocity uses the
   rotated index (sdf_collision) or the node (sdf_collision_no_arg). */
typedef struct orc_collider {
  const float *sdf, *normal;
  int32_t size[3], offset[3], center[3];
  int32_t from_rot, yshift, pad;
  double cos_t, sin_t, omega;
} orc_collider;

/* sdf_collision on grid velocities gv after orc_grid_op (3D, every node) */
void orc_collide_f32(const orc_mpm *s, const orc_collider *c, float *grid_v);
void orc_collide_f64(const orc_mpm *s, const orc_collider *c, double *grid_v);
/* the collider at node I (3D): 1 when inside, with the rotated normal and
   the collider velocity */
int orc_collider_probe_f32(const orc_mpm *s, const orc_collider *c, const int I[3], float nrm[3], float cv[3]);
int orc_collider_probe_f64(const orc_mpm *s, const orc_collider *c, const int I[3], double nrm[3], double cv[3]);

int orc_num_comps(const orc_mpm *s);

/* one phase each; grid arrays are caller-owned, p2g accumulates (caller zeroes) */
void orc_p2g_f32(const orc_mpm *s, const float *state, int64_t n, float *grid_m, float *grid_v);
void orc_grid_op_f32(const orc_mpm *s, const float *grid_m, float *grid_v);
void orc_g2p_f32(const orc_mpm *s, const float *state, const float *grid_v, float *out, int64_t n);
void orc_p2g_f64(const orc_mpm *s, const double *state, int64_t n, double *grid_m, double *grid_v);
void orc_grid_op_f64(const orc_mpm *s, const double *grid_m, double *grid_v);
void orc_g2p_f64(const orc_mpm *s, const double *state, const double *grid_v, double *out, int64_t n);

/* whole substeps on plain (unquantized) storage */
void orc_run_plain_f32(const orc_mpm *s, float *state, int64_t n, int nsteps,
                       float *grid_m, float *grid_v, int nthreads);
void orc_run_plain_f64(const orc_mpm *s, double *state, int64_t n, int nsteps,
                       double *grid_m, double *grid_v, int nthreads);
/* whole substeps on quantized storage (f32 compute): decode, substep, encode.
   Dithering noise of substep t is orc_dither_u(seed, step0 + t, p, c). */
void orc_run_quant_f32(const orc_mpm *s, const orc_comp *comps, int nwords,
                       uint64_t *words, int64_t n, int nsteps, uint64_t seed,
                       uint32_t step0, float *grid_m, float *grid_v,
                       float *scratch /* 2 * nc * n floats */, int nthreads);

/* the same with f64 compute: values decoded, stepped and encoded in double;
   scratch holds 2 * nc * n doubles. */
void orc_run_quant_f64(const orc_mpm *s, const orc_comp *comps, int nwords,
                       uint64_t *words, int64_t n, int nsteps, uint64_t seed,
                       uint32_t step0, double *grid_m, double *grid_v,
                       double *scratch, int nthreads);

/* Adjoint (reverse mode) of one plain f64 substep from state st:
   adj = dL/ds_t given adj_next = dL/ds_{t+1} (adj_ref.c); work holds
   orc_adj_work_doubles(s) doubles. */
int64_t orc_adj_work_doubles(const orc_mpm *s);
void orc_adj_substep_f64(const orc_mpm *s, const double *st, int64_t n, const double *adj_next,
                         double *adj, double *work);
/* the same with the SDF collider of the substep (NULL: none) after the grid
   update */
void orc_adj_substep_col_f64(const orc_mpm *s, const orc_collider *col, const double *st, int64_t n,
                             const double *adj_next, double *adj, double *work);
/* nsteps plain f64 substeps from state0 and their adjoint: returns the loss
   (kind 0: compute_ek; kind 1: sum(dir * s_T)), adj0 = dL/ds_0 and
   gsum[c] = sum over states t = 1..nsteps-1 and particles of (dL/ds_{t,p,c})^2. */
double orc_adjoint_run_f64(const orc_mpm *s, const double *state0, int64_t n, int nsteps, int loss_kind,
                           const double *dir, double *adj0, double *gsum);

/* compute_ek (mpm_solver.py:1043-1052), accumulated in double */
double orc_energy_f32(const orc_mpm *s, const float *state, int64_t n);
double orc_energy_f64(const orc_mpm *s, const double *state, int64_t n);
/* update_limits_* (mpm_solver.py:960-979): per-component max and min */
void orc_limits_f32(const orc_mpm *s, const float *state, int64_t n, float *mx, float *mn);

#ifdef __cplusplus
}
#endif
#endif
