/*
 * Oracle adjoint of the plain MLS-MPM substep (f64) -- TEST INFRASTRUCTURE ONLY.
 *
 * Reverse mode through one substep s_t -> s_{t+1} of mpm_solver.py:
 * p2g (:800-833), grid_op (:633-665), g2p (:835-858), the functions the
 * reference differentiates with Taichi's autodiff in
 * nlogn_loop_forward_grad (:1167-1198), with the SDF collider's
 * sdf_collision_no_arg (:1476-1504) when one is given, and the
 * error-contribution sums of
 * accu_grad / get_gradients (:1232-1302): g_c = sum over the intermediate
 * states t = 1..T-1 and particles p of (dL/ds_{t,p,c})^2, in f64.
 * Loss: compute_ek (:1043-1052) or a linear functional of the final state.
 *
 * The stencil base is treated as constant (piecewise), as in the
 * reference's autodiff; the quadratic B-spline weights are C1 across cell
 * boundaries, so the adjoint is the derivative almost everywhere.  Wall
 * conditions that zero a velocity component zero its adjoint.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_mpm.h"

static int64_t adj_nodes(const orc_mpm *s) {
  int64_t n = s->n_grid;
  return s->dim == 3 ? n * n * n : n * n;
}

static int adj_base(double xs, double inv_dx, int n) {
  double t = xs * inv_dx - 0.5;
  t = fmax(t, 0.0);
  t = fmin(t, (double)(n - 3));
  return (int)floor(t);
}

/* weights and their derivatives in fx */
static void adj_weights(double fx, double w[3], double dw[3]) {
  w[0] = 0.5 * (1.5 - fx) * (1.5 - fx);
  w[1] = 0.75 - (fx - 1.0) * (fx - 1.0);
  w[2] = 0.5 * (fx - 0.5) * (fx - 0.5);
  dw[0] = fx - 1.5;
  dw[1] = -2.0 * (fx - 1.0);
  dw[2] = fx - 0.5;
}

typedef struct {
  int base[3];
  double fx[3], w[3][3], dw[3][3];
} adj_stencil;

static void adj_stencil_of(const orc_mpm *s, const double *st, int64_t n, int64_t p, adj_stencil *S) {
  memset(S, 0, sizeof(*S));
  for (int a = 0; a < s->dim; a++) {
    double xs = st[(int64_t)a * n + p];
    S->base[a] = adj_base(xs, s->inv_dx, s->n_grid);
    S->fx[a] = xs * s->inv_dx - (double)S->base[a];
    adj_weights(S->fx[a], S->w[a], S->dw[a]);
  }
}

static int64_t adj_node(const orc_mpm *s, const adj_stencil *S, const int off[3]) {
  const int ng = s->n_grid;
  if (s->dim == 3)
    return ((int64_t)(S->base[0] + off[0]) * ng + (S->base[1] + off[1])) * ng + (S->base[2] + off[2]);
  return (int64_t)(S->base[0] + off[0]) * ng + (S->base[1] + off[1]);
}

/* weight of node off and its gradient in fx */
static double adj_weight(const orc_mpm *s, const adj_stencil *S, const int off[3], double grad[3]) {
  const int d = s->dim;
  double w = 1.0;
  for (int a = 0; a < d; a++) w *= S->w[a][off[a]];
  for (int c = 0; c < 3; c++) {
    grad[c] = 0.0;
    if (c >= d) continue;
    double g = S->dw[c][off[c]];
    for (int a = 0; a < d; a++)
      if (a != c) g *= S->w[a][off[a]];
    grad[c] = g;
  }
  return w;
}

/* stress (without the dt p_vol 4 inv_dx^2 factor) and its adjoint:
   sigma_bar -> (J_bar or F_bar) added into jf_bar */
static void adj_cauchy_back(const orc_mpm *s, const double *F /* J or F row major */,
                            const double *sb /* sigma_bar, d x d */, double *jf_bar) {
  const int d = s->dim;
  if (s->material == ORC_FLUID) {
    double tr = 0.0;
    for (int a = 0; a < d; a++) tr += sb[a * d + a];
    jf_bar[0] += s->la * tr;
    return;
  }
  if (d == 3) {
    /* sigma = mu (F F^T - I) + la log(J) on every entry (mpm_solver.py:1027-1029):
       F_bar = mu (sb + sb^T) F + la (sum sb) F^-T */
    double ssum = 0.0;
    for (int i = 0; i < 9; i++) ssum += sb[i];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double acc = 0.0;
        for (int j = 0; j < 3; j++) acc += (sb[a * 3 + j] + sb[j * 3 + a]) * F[j * 3 + b];
        jf_bar[a * 3 + b] += s->mu_0 * acc;
      }
    double cof[9];
    cof[0] = F[4] * F[8] - F[5] * F[7];
    cof[1] = F[5] * F[6] - F[3] * F[8];
    cof[2] = F[3] * F[7] - F[4] * F[6];
    cof[3] = F[2] * F[7] - F[1] * F[8];
    cof[4] = F[0] * F[8] - F[2] * F[6];
    cof[5] = F[1] * F[6] - F[0] * F[7];
    cof[6] = F[1] * F[5] - F[2] * F[4];
    cof[7] = F[2] * F[3] - F[0] * F[5];
    cof[8] = F[0] * F[4] - F[1] * F[3];
    const double J = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
    for (int i = 0; i < 9; i++) jf_bar[i] += s->la_0 * ssum * cof[i] / J;
    return;
  }
  /* 2D fixed-corotated: sigma = 2 mu (F - R) F^T + la J (J - 1) I */
  const double F00 = F[0], F01 = F[1], F10 = F[2], F11 = F[3];
  const double J = F00 * F11 - F01 * F10;
  const double x = F00 + F11, y = F10 - F01;
  const double r = sqrt(x * x + y * y);
  const double c = x / r, sn = y / r;
  const double R[4] = {c, -sn, sn, c};
  const double twomu = 2.0 * s->mu_0;
  double A[4];
  for (int i = 0; i < 4; i++) A[i] = twomu * (F[i] - R[i]);
  double Fb[4] = {0, 0, 0, 0};
  /* P = A F^T: A_bar = sigma_bar F, F_bar += sigma_bar^T A */
  double Ab[4];
  for (int i = 0; i < 2; i++)
    for (int k = 0; k < 2; k++) {
      Ab[i * 2 + k] = sb[i * 2 + 0] * F[0 * 2 + k] + sb[i * 2 + 1] * F[1 * 2 + k];
      Fb[i * 2 + k] += sb[0 * 2 + i] * A[0 * 2 + k] + sb[1 * 2 + i] * A[1 * 2 + k];
    }
  /* A = 2 mu (F - R) */
  double Rb[4];
  for (int i = 0; i < 4; i++) {
    Fb[i] += twomu * Ab[i];
    Rb[i] = -twomu * Ab[i];
  }
  /* R = [[c, -s], [s, c]], c = x / r, s = y / r */
  const double cb = Rb[0] + Rb[3], sb_ = Rb[2] - Rb[1];
  const double r3 = r * r * r;
  const double xb = cb * (y * y / r3) - sb_ * (x * y / r3);
  const double yb = -cb * (x * y / r3) + sb_ * (x * x / r3);
  Fb[0] += xb;
  Fb[3] += xb;
  Fb[2] += yb;
  Fb[1] -= yb;
  /* la J (J - 1) on the diagonal */
  const double Jb = s->la_0 * (2.0 * J - 1.0) * (sb[0] + sb[3]);
  Fb[0] += Jb * F11;
  Fb[3] += Jb * F00;
  Fb[1] -= Jb * F10;
  Fb[2] -= Jb * F01;
  for (int i = 0; i < 4; i++) jf_bar[i] += Fb[i];
}

void orc_adj_substep_col_f64(const orc_mpm *s, const orc_collider *col, const double *st, int64_t n,
                             const double *adj_next, double *adj, double *work) {
  const int d = s->dim, nc = orc_num_comps(s);
  const int cV = d, cC = 2 * d, cJF = 2 * d + d * d, njf = nc - cJF;
  const int64_t nodes = adj_nodes(s);
  double *gm = work, *gmv = work + nodes, *gv = gmv + (int64_t)d * nodes;
  double *gbar = gv + (int64_t)d * nodes, *mbar = gbar + (int64_t)d * nodes;
  const double dx = s->dx, inv_dx = s->inv_dx, dt = s->dt, pm = s->p_mass, c4 = 4.0 * s->inv_dx;
  const double stress_k = -(s->dt * s->p_vol * 4 * s->inv_dx * s->inv_dx);
  const int nk = d == 3 ? 3 : 1;

  /* forward recompute: P2G, grid update */
  memset(gm, 0, sizeof(double) * (size_t)nodes * (size_t)(1 + d));
  orc_p2g_f64(s, st, n, gm, gmv);
  memcpy(gv, gmv, sizeof(double) * (size_t)nodes * (size_t)d);
  orc_grid_op_f64(s, gm, gv);
  if (col && d == 3) orc_collide_f64(s, col, gv);
  memset(gbar, 0, sizeof(double) * (size_t)nodes * (size_t)(d + 1));
  memset(adj, 0, sizeof(double) * (size_t)nc * (size_t)n);

  /* G2P backward: particle adjoints of x and J/F, grid velocity adjoint */
  for (int64_t p = 0; p < n; p++) {
    adj_stencil S;
    adj_stencil_of(s, st, n, p, &S);
    double nv[3] = {0, 0, 0}, nC[9] = {0};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < nk; k++) {
          int off[3] = {i, j, k};
          double gw[3];
          double w = adj_weight(s, &S, off, gw);
          int64_t node = adj_node(s, &S, off);
          for (int a = 0; a < d; a++) {
            double g = gv[(int64_t)a * nodes + node];
            nv[a] += w * g;
            for (int b = 0; b < d; b++) nC[a * d + b] += c4 * w * g * ((double)off[b] - S.fx[b]);
          }
        }
    double xb_next[3], vbar[3], Cbar[9];
    for (int a = 0; a < d; a++) {
      xb_next[a] = adj_next[(int64_t)a * n + p];
      vbar[a] = adj_next[(int64_t)(cV + a) * n + p] + dt * xb_next[a];
    }
    for (int i = 0; i < d * d; i++) Cbar[i] = adj_next[(int64_t)(cC + i) * n + p];
    double jf[9], jfb[9] = {0};
    for (int i = 0; i < njf; i++) jf[i] = st[(int64_t)(cJF + i) * n + p];
    if (s->material == ORC_FLUID) {
      double tr = 0.0;
      for (int a = 0; a < d; a++) tr += nC[a * d + a];
      double Jb = adj_next[(int64_t)cJF * n + p];
      jfb[0] = Jb * (1.0 + dt * tr);
      for (int a = 0; a < d; a++) Cbar[a * d + a] += Jb * dt * jf[0];
    } else {
      /* F' = (I + dt C') F */
      double Fbn[9], M[9];
      for (int i = 0; i < d * d; i++) Fbn[i] = adj_next[(int64_t)(cJF + i) * n + p];
      for (int a = 0; a < d; a++)
        for (int b = 0; b < d; b++) M[a * d + b] = (a == b ? 1.0 : 0.0) + dt * nC[a * d + b];
      for (int a = 0; a < d; a++)
        for (int b = 0; b < d; b++) {
          double acc = 0.0, acc2 = 0.0;
          for (int k = 0; k < d; k++) {
            acc += M[k * d + a] * Fbn[k * d + b];   /* M^T F_bar' */
            acc2 += Fbn[a * d + k] * jf[b * d + k]; /* F_bar' F^T */
          }
          jfb[a * d + b] = acc;
          Cbar[a * d + b] += dt * acc2;
        }
    }
    double fxb[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < nk; k++) {
          int off[3] = {i, j, k};
          double gw[3];
          double w = adj_weight(s, &S, off, gw);
          int64_t node = adj_node(s, &S, off);
          double g[3] = {0, 0, 0}, u[3] = {0, 0, 0};
          for (int a = 0; a < d; a++) g[a] = gv[(int64_t)a * nodes + node];
          for (int a = 0; a < d; a++) {
            u[a] = vbar[a];
            for (int b = 0; b < d; b++) u[a] += c4 * Cbar[a * d + b] * ((double)off[b] - S.fx[b]);
            gbar[(int64_t)a * nodes + node] += w * u[a];
          }
          double gu = 0.0;
          for (int a = 0; a < d; a++) gu += g[a] * u[a];
          for (int c = 0; c < d; c++) {
            double t = 0.0;
            for (int a = 0; a < d; a++) t += Cbar[a * d + c] * g[a];
            fxb[c] += gw[c] * gu - c4 * w * t;
          }
        }
    for (int a = 0; a < d; a++) adj[(int64_t)a * n + p] = xb_next[a] + inv_dx * fxb[a];
    for (int i = 0; i < njf; i++) adj[(int64_t)(cJF + i) * n + p] = jfb[i];
  }

  /* grid update backward (walls per axis, mpm_solver.py:97-105), after the
     collider's: sdf_collision_no_arg (:1476-1504) moves a node inside the
     collider whose velocity after the walls approaches it (v_rel . n < 0)
     to v - (v_rel . n) n, Jacobian I - n n^T */
  int lb[3], ub[3];
  for (int a = 0; a < 3; a++) {
    lb[a] = s->axis_bounds && a < d ? s->lb[a] : s->bound;
    ub[a] = s->axis_bounds && a < d ? s->ub[a] : s->n_grid - s->bound;
  }
  const int ng = s->n_grid;
  const double gdt = s->dt * s->gravity;
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
    const double inv_m = 1.0 / (gm[node] + 1e-10);
    double v[3];
    for (int a = 0; a < d; a++) v[a] = inv_m * gmv[(int64_t)a * nodes + node];
    v[1] -= gdt;
    double gb[3];
    for (int a = 0; a < d; a++) gb[a] = gbar[(int64_t)a * nodes + node];
    double vw[3] = {v[0], v[1], d == 3 ? v[2] : 0.0};
    int wall = 0;
    if (s->boundary == ORC_STICKY) {
      for (int a = 0; a < d; a++)
        if (I[a] < lb[a] || I[a] > ub[a]) wall = 1;
      if (wall)
        for (int a = 0; a < d; a++) vw[a] = 0.0;
    } else {
      for (int a = 0; a < d; a++) {
        if (I[a] < lb[a] && v[a] < 0) vw[a] = 0.0;
        if (I[a] > ub[a] && v[a] > 0) vw[a] = 0.0;
      }
    }
    double nrm[3], cv[3];
    if (col && d == 3 && orc_collider_probe_f64(s, col, I, nrm, cv)) {
      const double dn = (vw[0] - cv[0]) * nrm[0] + (vw[1] - cv[1]) * nrm[1] + (vw[2] - cv[2]) * nrm[2];
      if (dn < 0.0) {
        const double g = gb[0] * nrm[0] + gb[1] * nrm[1] + gb[2] * nrm[2];
        for (int a = 0; a < 3; a++) gb[a] -= g * nrm[a];
      }
    }
    if (s->boundary == ORC_STICKY) {
      if (wall)
        for (int a = 0; a < d; a++) gb[a] = 0.0;
    } else {
      for (int a = 0; a < d; a++) {
        if (I[a] < lb[a] && v[a] < 0) gb[a] = 0.0;
        if (I[a] > ub[a] && v[a] > 0) gb[a] = 0.0;
      }
    }
    double mb = 0.0;
    for (int a = 0; a < d; a++) {
      mb -= gb[a] * gmv[(int64_t)a * nodes + node] * inv_m * inv_m;
      gbar[(int64_t)a * nodes + node] = gb[a] * inv_m; /* now d L / d (m v) */
    }
    mbar[node] = mb;
  }

  /* P2G backward */
  for (int64_t p = 0; p < n; p++) {
    adj_stencil S;
    adj_stencil_of(s, st, n, p, &S);
    double vv[3] = {0, 0, 0}, C[9], jf[9];
    for (int a = 0; a < d; a++) vv[a] = st[(int64_t)(cV + a) * n + p];
    for (int i = 0; i < d * d; i++) C[i] = st[(int64_t)(cC + i) * n + p];
    for (int i = 0; i < njf; i++) jf[i] = st[(int64_t)(cJF + i) * n + p];
    /* affine = stress_k sigma + m C */
    double sig[9] = {0};
    if (s->material == ORC_FLUID) {
      for (int a = 0; a < d; a++) sig[a * d + a] = s->la * (jf[0] - 1.0);
    } else if (d == 3) {
      const double J = jf[0] * (jf[4] * jf[8] - jf[7] * jf[5]) - jf[3] * (jf[1] * jf[8] - jf[7] * jf[2]) +
                       jf[6] * (jf[1] * jf[5] - jf[4] * jf[2]);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double fft = jf[i * 3 + 0] * jf[j * 3 + 0] + jf[i * 3 + 1] * jf[j * 3 + 1] + jf[i * 3 + 2] * jf[j * 3 + 2];
          sig[i * 3 + j] = s->mu_0 * (fft - (i == j ? 1.0 : 0.0)) + s->la_0 * log(J);
        }
    } else {
      const double F00 = jf[0], F01 = jf[1], F10 = jf[2], F11 = jf[3];
      const double J = F00 * F11 - F01 * F10;
      const double x = F00 + F11, y = F10 - F01, r = sqrt(x * x + y * y);
      const double R[4] = {x / r, -y / r, y / r, x / r};
      double A[4];
      for (int i = 0; i < 4; i++) A[i] = 2.0 * s->mu_0 * (jf[i] - R[i]);
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) sig[i * 2 + j] = A[i * 2 + 0] * jf[j * 2 + 0] + A[i * 2 + 1] * jf[j * 2 + 1];
      sig[0] += s->la_0 * J * (J - 1.0);
      sig[3] += s->la_0 * J * (J - 1.0);
    }
    double aff[9];
    for (int i = 0; i < d * d; i++) aff[i] = stress_k * sig[i] + pm * C[i];
    double vb[3] = {0, 0, 0}, affb[9] = {0}, fxb[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < nk; k++) {
          int off[3] = {i, j, k};
          double gw[3];
          double w = adj_weight(s, &S, off, gw);
          int64_t node = adj_node(s, &S, off);
          double dp[3] = {0, 0, 0}, mvb[3] = {0, 0, 0};
          for (int a = 0; a < d; a++) {
            dp[a] = ((double)off[a] - S.fx[a]) * dx;
            mvb[a] = gbar[(int64_t)a * nodes + node];
          }
          const double mb = mbar[node];
          double contrib = mb * pm;
          for (int a = 0; a < d; a++) {
            double q = pm * vv[a];
            for (int b = 0; b < d; b++) q += aff[a * d + b] * dp[b];
            contrib += mvb[a] * q;
            vb[a] += pm * w * mvb[a];
            for (int b = 0; b < d; b++) affb[a * d + b] += w * mvb[a] * dp[b];
          }
          for (int c = 0; c < d; c++) {
            double t = 0.0;
            for (int a = 0; a < d; a++) t += aff[a * d + c] * mvb[a];
            fxb[c] += gw[c] * contrib - dx * w * t;
          }
        }
    for (int a = 0; a < d; a++) {
      adj[(int64_t)a * n + p] += inv_dx * fxb[a];
      adj[(int64_t)(cV + a) * n + p] += vb[a];
    }
    for (int i = 0; i < d * d; i++) adj[(int64_t)(cC + i) * n + p] += pm * affb[i];
    double sb[9], jfb[9] = {0};
    for (int i = 0; i < d * d; i++) sb[i] = stress_k * affb[i];
    adj_cauchy_back(s, jf, sb, jfb);
    for (int i = 0; i < njf; i++) adj[(int64_t)(cJF + i) * n + p] += jfb[i];
  }
}

void orc_adj_substep_f64(const orc_mpm *s, const double *st, int64_t n, const double *adj_next, double *adj,
                         double *work) {
  orc_adj_substep_col_f64(s, NULL, st, n, adj_next, adj, work);
}

int64_t orc_adj_work_doubles(const orc_mpm *s) { return adj_nodes(s) * (int64_t)(2 * s->dim + 2 + s->dim); }

/* dL/ds_T of the loss: kind 0 compute_ek (2D sum |v|^2, 3D sum 1/2 m |v|^2 + g m y),
   kind 1 sum dir * s_T.  Returns the loss. */
static double adj_loss(const orc_mpm *s, const double *sT, int64_t n, int kind, const double *dir, double *adjT) {
  const int d = s->dim, nc = orc_num_comps(s);
  memset(adjT, 0, sizeof(double) * (size_t)nc * (size_t)n);
  double L = 0.0;
  if (kind == 1) {
    for (int64_t i = 0; i < (int64_t)nc * n; i++) {
      adjT[i] = dir[i];
      L += dir[i] * sT[i];
    }
    return L;
  }
  for (int64_t p = 0; p < n; p++) {
    for (int a = 0; a < d; a++) {
      double v = sT[(int64_t)(d + a) * n + p];
      if (d == 3) {
        L += 0.5 * s->p_mass * v * v;
        adjT[(int64_t)(d + a) * n + p] = s->p_mass * v;
      } else {
        L += v * v;
        adjT[(int64_t)(d + a) * n + p] = 2.0 * v;
      }
    }
    if (d == 3) {
      L += s->gravity * s->p_mass * sT[(int64_t)1 * n + p];
      adjT[(int64_t)1 * n + p] = s->gravity * s->p_mass;
    }
  }
  return L;
}

double orc_adjoint_run_f64(const orc_mpm *s, const double *state0, int64_t n, int nsteps, int loss_kind,
                           const double *dir, double *adj0, double *gsum) {
  const int nc = orc_num_comps(s);
  const int64_t sz = (int64_t)nc * n;
  const int64_t nodes = adj_nodes(s);
  double *states = (double *)malloc(sizeof(double) * (size_t)sz * (size_t)(nsteps + 1));
  double *gm = (double *)malloc(sizeof(double) * (size_t)nodes * (size_t)(1 + s->dim));
  double *work = (double *)malloc(sizeof(double) * (size_t)orc_adj_work_doubles(s));
  double *a_next = (double *)malloc(sizeof(double) * (size_t)sz);
  memcpy(states, state0, sizeof(double) * (size_t)sz);
  for (int t = 0; t < nsteps; t++) {
    memcpy(states + (t + 1) * sz, states + t * sz, sizeof(double) * (size_t)sz);
    orc_run_plain_f64(s, states + (t + 1) * sz, n, 1, gm, gm + nodes, 1);
  }
  const double L = adj_loss(s, states + (int64_t)nsteps * sz, n, loss_kind, dir, a_next);
  for (int c = 0; c < nc; c++) gsum[c] = 0.0;
  for (int t = nsteps - 1; t >= 0; t--) {
    orc_adj_substep_f64(s, states + (int64_t)t * sz, n, a_next, adj0, work);
    if (t >= 1)
      for (int c = 0; c < nc; c++)
        for (int64_t p = 0; p < n; p++) gsum[c] += adj0[(int64_t)c * n + p] * adj0[(int64_t)c * n + p];
    memcpy(a_next, adj0, sizeof(double) * (size_t)sz);
  }
  free(states);
  free(gm);
  free(work);
  free(a_next);
  return L;
}
