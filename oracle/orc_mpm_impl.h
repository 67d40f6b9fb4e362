/*
 * Body of the oracle MPM phases, included by mpm_ref.c once per compute
 * type with REAL (float/double) and SFX (f32/f64) defined.  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Expressions keep the reference's operand order so that each rounding
 * step is the reference's (unfused):
 *   base/fx/w            mpm_solver.py:599-606, 676-678
 *   stress/affine        mpm_solver.py:610-613, compute_cauchy :1012-1031
 *   P2G scatter          mpm_solver.py:615-622
 *   grid update          mpm_solver.py:633-665 (friction off)
 *   G2P gather + update  mpm_solver.py:682-694, 702, 750-758
 * Deviation: the stencil base is clamped to [0, n_grid-3] so a particle that
 * left the domain cannot index outside the dense grid (the reference's
 * sparse grid spans 4096 cells and would accept it).
 */
#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

static int FN(base_of)(REAL xs, REAL inv_dx, int n) {
  REAL t = xs * inv_dx;               /* x * inv_dx */
  t = t - (REAL)0.5;                  /* - 0.5 */
  /* clamp before floor (NaN -> 0): identical to floor-then-clamp for finite x */
  t = FMAX(t, (REAL)0);
  t = FMIN(t, (REAL)(n - 3));
  return (int)FLOOR(t);               /* floor(...).cast(int) */
}

static void FN(weights)(REAL fx, REAL w[3]) {
  REAL a = (REAL)1.5 - fx;
  w[0] = (REAL)0.5 * (a * a);
  REAL b = fx - (REAL)1;
  w[1] = (REAL)0.75 - b * b;
  REAL c = fx - (REAL)0.5;
  w[2] = (REAL)0.5 * (c * c);
}

/* compute_cauchy (mpm_solver.py:1012-1031) times the stress coefficient,
   plus p_mass * C (:610-613).  aff is row major dim x dim. */
static void FN(affine)(const orc_mpm *s, const REAL *st, int64_t n, int64_t p,
                       REAL *aff) {
  const int d = s->dim;
  const int cC = 2 * d, cJF = 2 * d + d * d;
  REAL cauchy[9];
  for (int i = 0; i < d * d; i++) cauchy[i] = (REAL)0;
  if (s->material == ORC_FLUID) {
    REAL J = st[(int64_t)cJF * n + p];
    REAL sig = (REAL)s->la * (J - (REAL)1.0);
    for (int i = 0; i < d; i++) cauchy[i * d + i] = sig;
  } else if (d == 3) {
    /* 3D: mu (F F^T - I) + la log(J), the scalar added to every entry
       (mpm_solver.py:1027-1029); J by Matrix.determinant's cofactor
       expansion (TS/python/taichi/lang/matrix.py:625-628) */
    REAL F[9];
    for (int i = 0; i < 9; i++) F[i] = st[(int64_t)(cJF + i) * n + p];
    REAL J = F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
             F[6] * (F[1] * F[5] - F[4] * F[2]);
    REAL mu = (REAL)s->mu_0, lg = (REAL)s->la_0 * LOG(J);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        REAL fft = F[i * 3 + 0] * F[j * 3 + 0];
        fft = fft + F[i * 3 + 1] * F[j * 3 + 1];
        fft = fft + F[i * 3 + 2] * F[j * 3 + 2];
        cauchy[i * 3 + j] = mu * (fft - (i == j ? (REAL)1 : (REAL)0)) + lg;
      }
  } else {
    /* 2D fixed-corotated */
    REAL F00 = st[(int64_t)(cJF + 0) * n + p], F01 = st[(int64_t)(cJF + 1) * n + p];
    REAL F10 = st[(int64_t)(cJF + 2) * n + p], F11 = st[(int64_t)(cJF + 3) * n + p];
    REAL J = F00 * F11 - F01 * F10;                 /* Matrix.determinant 2x2 */
    /* polar_decompose2d, linalg_impl.py:9-26 */
    REAL x = F00 + F11, y = F10 - F01;
    REAL sc = (REAL)1.0 / SQRT(x * x + y * y);
    REAL c = x * sc, sn = y * sc;
    REAL R00 = c, R01 = -sn, R10 = sn, R11 = c;
    REAL twomu = (REAL)(2 * s->mu_0);
    REAL A00 = twomu * (F00 - R00), A01 = twomu * (F01 - R01);
    REAL A10 = twomu * (F10 - R10), A11 = twomu * (F11 - R11);
    /* (2 mu (F - R)) @ F^T */
    REAL P00 = A00 * F00 + A01 * F01, P01 = A00 * F10 + A01 * F11;
    REAL P10 = A10 * F00 + A11 * F01, P11 = A10 * F10 + A11 * F11;
    /* + ((I * la) * J) * (J - 1) */
    REAL dg = ((REAL)s->la_0 * J) * (J - (REAL)1);
    REAL og = (((REAL)0 * (REAL)s->la_0) * J) * (J - (REAL)1);
    cauchy[0] = P00 + dg;
    cauchy[1] = P01 + og;
    cauchy[2] = P10 + og;
    cauchy[3] = P11 + dg;
  }
  REAL k = (REAL)(-(s->dt * s->p_vol * 4 * s->inv_dx * s->inv_dx));
  REAL pm = (REAL)s->p_mass;
  for (int i = 0; i < d * d; i++)
    aff[i] = k * cauchy[i] + pm * st[(int64_t)(cC + i) * n + p];
}

/* P2G of particles [lo, hi) of a state with plane stride n */
static void FN(p2g_span)(const orc_mpm *s, const REAL *st, int64_t n, int64_t lo, int64_t hi,
                         REAL *gm, REAL *gv) {
  const int d = s->dim, ng = s->n_grid;
  const int64_t nodes = d == 3 ? (int64_t)ng * ng * ng : (int64_t)ng * ng;
  const REAL inv_dx = (REAL)s->inv_dx, dx = (REAL)s->dx, pm = (REAL)s->p_mass;
  for (int64_t p = lo; p < hi; p++) {
    int base[3] = {0, 0, 0};
    REAL fx[3] = {0, 0, 0}, w[3][3];
    for (int a = 0; a < d; a++) {
      REAL xs = st[(int64_t)a * n + p];
      base[a] = FN(base_of)(xs, inv_dx, ng);
      fx[a] = xs * inv_dx - (REAL)base[a];
      FN(weights)(fx[a], w[a]);
    }
    REAL aff[9];
    FN(affine)(s, st, n, p, aff);
    REAL mv[3] = {0, 0, 0};
    for (int a = 0; a < d; a++) mv[a] = pm * st[(int64_t)(d + a) * n + p];
    const int nk = d == 3 ? 3 : 1;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < nk; k++) {
          int off[3] = {i, j, k};
          REAL dpos[3];
          for (int a = 0; a < d; a++) dpos[a] = ((REAL)off[a] - fx[a]) * dx;
          REAL weight = (REAL)1.0;
          for (int a = 0; a < d; a++) weight *= w[a][off[a]];
          int64_t node = d == 3 ? ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k)
                                : (int64_t)(base[0] + i) * ng + (base[1] + j);
          for (int a = 0; a < d; a++) {
            REAL ad = aff[a * d + 0] * dpos[0];
            for (int b = 1; b < d; b++) ad = ad + aff[a * d + b] * dpos[b];
            gv[(int64_t)a * nodes + node] += weight * (mv[a] + ad);
          }
          gm[node] += weight * pm;
        }
  }
}

void FN(orc_p2g)(const orc_mpm *s, const REAL *st, int64_t n, REAL *gm, REAL *gv) {
  FN(p2g_span)(s, st, n, 0, n, gm, gv);
}

/* walls per axis: lb = bound, ub = n_grid - bound, or MPMConfig.lb / ub
   (mpm_solver.py:97-105) */
static void FN(walls)(const orc_mpm *s, int lb[3], int ub[3]) {
  for (int a = 0; a < 3; a++) {
    lb[a] = s->axis_bounds && a < s->dim ? s->lb[a] : s->bound;
    ub[a] = s->axis_bounds && a < s->dim ? s->ub[a] : s->n_grid - s->bound;
  }
}

void FN(orc_grid_op)(const orc_mpm *s, const REAL *gm, REAL *gv) {
  const int d = s->dim, ng = s->n_grid;
  const int64_t nodes = d == 3 ? (int64_t)ng * ng * ng : (int64_t)ng * ng;
  const REAL g = (REAL)(s->dt * s->gravity);
  int lb[3], ub[3];
  FN(walls)(s, lb, ub);
  for (int64_t node = 0; node < nodes; node++) {
    int I[3];
    if (d == 3) {
      I[0] = (int)(node / ((int64_t)ng * ng));
      I[1] = (int)((node / ng) % ng);
      I[2] = (int)(node % ng);
    } else {
      I[0] = (int)(node / ng);
      I[1] = (int)(node % ng);
      I[2] = 0;
    }
    REAL inv_m = (REAL)1 / (gm[node] + (REAL)1e-10);
    REAL v[3];
    for (int a = 0; a < d; a++) v[a] = inv_m * gv[(int64_t)a * nodes + node];
    v[1] = v[1] - g;
    if (s->boundary == ORC_STICKY) {
      int wall = 0;
      for (int a = 0; a < d; a++)
        if (I[a] < lb[a] || I[a] > ub[a]) wall = 1;
      if (wall)
        for (int a = 0; a < d; a++) v[a] = (REAL)0;
    } else {
      for (int a = 0; a < d; a++) {
        if (I[a] < lb[a] && v[a] < 0) v[a] = (REAL)0;
        if (I[a] > ub[a] && v[a] > 0) v[a] = (REAL)0;
      }
    }
    for (int a = 0; a < d; a++) gv[(int64_t)a * nodes + node] = v[a];
  }
}

/* The collider at node I: rot_scene (mpm_solver.py:1524-1545), the node's
   offset from the centre rotated by -theta, shifted down by the height,
   truncated and clamped to the walls; sdf (:1447-1452): inside where the
   stored value < -1e-5 (outside the volume the reference's field holds 0).
   Returns 1 when inside, with the normal rotated by theta (nrm) and the
   collider velocity omega e_y x r (cv). */
int FN(orc_collider_probe)(const orc_mpm *s, const orc_collider *c, const int I[3], REAL nrm[3], REAL cv[3]) {
  int lb[3], ub[3];
  FN(walls)(s, lb, ub);
  const REAL ct = (REAL)c->cos_t, st = (REAL)c->sin_t, om = (REAL)c->omega, dx = (REAL)s->dx;
  int idx[3] = {I[0], I[1], I[2]};
  if (idx[0] >= 2 * c->center[0]) idx[0] -= 2 * c->center[0];
  if (idx[2] >= 2 * c->center[2]) idx[2] -= 2 * c->center[2];
  const REAL o0 = (REAL)(idx[0] - c->center[0]), o2 = (REAL)(idx[2] - c->center[2]);
  /* cos(-theta) o0 + sin(-theta) o2 + C0 + 0.5; -sin(-theta) o0 + cos(-theta) o2 + C2 + 0.5 */
  const REAL x = ct * o0 + (-st) * o2 + (REAL)c->center[0] + (REAL)0.5;
  const REAL z = -(-st) * o0 + ct * o2 + (REAL)c->center[2] + (REAL)0.5;
  int r[3] = {(int)x, I[1] - c->yshift, (int)z};
  for (int a = 0; a < 3; a++) {
    if (r[a] > ub[a]) r[a] = ub[a];
    if (r[a] < lb[a]) r[a] = lb[a];
  }
  int q[3];
  for (int a = 0; a < 3; a++) {
    q[a] = r[a] - c->offset[a];
    if (q[a] < 0 || q[a] >= c->size[a]) return 0;
  }
  const int64_t e = ((int64_t)q[0] * c->size[1] + q[1]) * c->size[2] + q[2];
  if (!(c->sdf[e] < -1e-5f)) return 0;
  const REAL g0 = (REAL)c->normal[3 * e], g1 = (REAL)c->normal[3 * e + 1], g2 = (REAL)c->normal[3 * e + 2];
  nrm[0] = ct * g0 + st * g2;
  nrm[1] = g1;
  nrm[2] = -st * g0 + ct * g2;
  const int p0 = c->from_rot ? r[0] : I[0], p2 = c->from_rot ? r[2] : I[2];
  const REAL rr[3] = {(REAL)(p0 - c->center[0]) * dx, (REAL)0, (REAL)(p2 - c->center[2]) * dx};
  /* omega x r with omega = (0, om, 0) */
  cv[0] = om * rr[2] - (REAL)0 * rr[1];
  cv[1] = (REAL)0 * rr[0] - (REAL)0 * rr[2];
  cv[2] = (REAL)0 * rr[1] - om * rr[0];
  return 1;
}

/* sdf_collision (mpm_solver.py:1459-1521) over every node of the dense 3D
   grid: v = cv + (v_rel - min(v_rel . n, 0) n), v_rel = v - cv. */
void FN(orc_collide)(const orc_mpm *s, const orc_collider *c, REAL *gv) {
  const int ng = s->n_grid;
  const int64_t nodes = (int64_t)ng * ng * ng;
  for (int64_t node = 0; node < nodes; node++) {
    const int I[3] = {(int)(node / ((int64_t)ng * ng)), (int)((node / ng) % ng), (int)(node % ng)};
    REAL nrm[3], cv[3];
    if (!FN(orc_collider_probe)(s, c, I, nrm, cv)) continue;
    REAL vr[3];
    for (int a = 0; a < 3; a++) vr[a] = gv[(int64_t)a * nodes + node] - cv[a];
    const REAL dn = vr[0] * nrm[0] + vr[1] * nrm[1] + vr[2] * nrm[2];
    const REAL m = FMIN(dn, (REAL)0);
    for (int a = 0; a < 3; a++) gv[(int64_t)a * nodes + node] = cv[a] + (vr[a] - m * nrm[a]);
  }
}

/* G2P of particles [lo, hi) of a state with plane stride n */
static void FN(g2p_span)(const orc_mpm *s, const REAL *st, const REAL *gv, REAL *out, int64_t n,
                         int64_t lo, int64_t hi) {
  const int d = s->dim, ng = s->n_grid;
  const int64_t nodes = d == 3 ? (int64_t)ng * ng * ng : (int64_t)ng * ng;
  const REAL inv_dx = (REAL)s->inv_dx, dt = (REAL)s->dt;
  const REAL c4 = (REAL)(4 * s->inv_dx);
  const int cC = 2 * d, cJF = 2 * d + d * d;
  for (int64_t p = lo; p < hi; p++) {
    int base[3] = {0, 0, 0};
    REAL fx[3] = {0, 0, 0}, w[3][3];
    for (int a = 0; a < d; a++) {
      REAL xs = st[(int64_t)a * n + p];
      base[a] = FN(base_of)(xs, inv_dx, ng);
      fx[a] = xs * inv_dx - (REAL)base[a];
      FN(weights)(fx[a], w[a]);
    }
    REAL nv[3] = {0, 0, 0}, nC[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int nk = d == 3 ? 3 : 1;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < nk; k++) {
          int off[3] = {i, j, k};
          REAL dpos[3];
          for (int a = 0; a < d; a++) dpos[a] = (REAL)off[a] - fx[a];
          int64_t node = d == 3 ? ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k)
                                : (int64_t)(base[0] + i) * ng + (base[1] + j);
          REAL weight = (REAL)1.0;
          for (int a = 0; a < d; a++) weight *= w[a][off[a]];
          REAL g[3];
          for (int a = 0; a < d; a++) g[a] = gv[(int64_t)a * nodes + node];
          for (int a = 0; a < d; a++) nv[a] = nv[a] + weight * g[a];
          REAL cw = c4 * weight;
          for (int a = 0; a < d; a++)
            for (int b = 0; b < d; b++) nC[a * d + b] = nC[a * d + b] + cw * (g[a] * dpos[b]);
        }
    int lb[3], ub[3];
    FN(walls)(s, lb, ub);
    for (int a = 0; a < d; a++) {
      REAL x = st[(int64_t)a * n + p] + dt * nv[a];
      /* new_x[k] = max(min(new_x[k], ub[k] * dx), lb[k] * dx), the bounds
         folded from Python doubles */
      if (s->bound_clip) x = FMAX(FMIN(x, (REAL)(ub[a] * s->dx)), (REAL)(lb[a] * s->dx));
      out[(int64_t)a * n + p] = x;
      out[(int64_t)(d + a) * n + p] = nv[a];
    }
    for (int i = 0; i < d * d; i++) out[(int64_t)(cC + i) * n + p] = nC[i];
    if (s->material == ORC_FLUID) {
      REAL tr = nC[0];
      for (int a = 1; a < d; a++) tr = tr + nC[a * d + a];
      out[(int64_t)cJF * n + p] = ((REAL)1.0 + dt * tr) * st[(int64_t)cJF * n + p];
    } else {
      /* (I + dt C) @ F */
      REAL M[9], F[9];
      for (int a = 0; a < d; a++)
        for (int b = 0; b < d; b++) {
          M[a * d + b] = (a == b ? (REAL)1 : (REAL)0) + dt * nC[a * d + b];
          F[a * d + b] = st[(int64_t)(cJF + a * d + b) * n + p];
        }
      for (int a = 0; a < d; a++)
        for (int b = 0; b < d; b++) {
          REAL acc = M[a * d + 0] * F[0 * d + b];
          for (int k = 1; k < d; k++) acc = acc + M[a * d + k] * F[k * d + b];
          out[(int64_t)(cJF + a * d + b) * n + p] = acc;
        }
    }
  }
}

void FN(orc_g2p)(const orc_mpm *s, const REAL *st, const REAL *gv, REAL *out, int64_t n) {
  FN(g2p_span)(s, st, gv, out, n, 0, n);
}

double FN(orc_energy)(const orc_mpm *s, const REAL *st, int64_t n) {
  const int d = s->dim;
  double e = 0.0;
  for (int64_t p = 0; p < n; p++) {
    REAL v2 = (REAL)0;
    for (int a = 0; a < d; a++) {
      REAL v = st[(int64_t)(d + a) * n + p];
      v2 = v2 + v * v;
    }
    if (d == 3) {
      e += (double)((REAL)(0.5 * s->p_mass) * v2);
      e += (double)((REAL)(s->gravity * s->p_mass) * st[(int64_t)1 * n + p]);
    } else {
      e += (double)v2;
    }
  }
  return e;
}

static void FN(zero_grid)(const orc_mpm *s, REAL *gm, REAL *gv) {
  const int d = s->dim, ng = s->n_grid;
  const int64_t nodes = d == 3 ? (int64_t)ng * ng * ng : (int64_t)ng * ng;
  memset(gm, 0, sizeof(REAL) * (size_t)nodes);
  memset(gv, 0, sizeof(REAL) * (size_t)nodes * (size_t)d);
}

#undef FN
#undef CAT
#undef CAT_
