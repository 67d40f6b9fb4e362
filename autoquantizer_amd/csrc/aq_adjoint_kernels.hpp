// Adjoint (reverse mode) of the plain MLS-MPM substep on gfx950.
//
// The reference differentiates p2g / grid_op / g2p with Taichi's autodiff
// (mpm_exps/engine/mpm_solver.py:800-858, nlogn_loop_forward_grad
// :1167-1188) and accumulates the squared state adjoints per component
// (accu_grad :1232-1271).  Here the three adjoint phases are hand-written:
//   k_adj_g2p   per particle: adjoints of x and J/F through G2P, and the
//               grid-velocity adjoint scattered with float atomics;
//   k_adj_grid  per node: adjoint of the grid update, (m v, m) adjoints;
//   k_adj_p2g   per particle: adjoints of x, v, C, J/F through P2G (gathered),
//               plus the squared adjoints summed per component in f64.
// The stencil base is piecewise constant; quadratic B-spline weights are C1,
// so this is the derivative almost everywhere.  Wall conditions that zero a
// velocity component zero its adjoint.
#pragma once

#include "aq_gather.hpp"  // collider_probe
#include "aq_mpm_kernels.hpp"

namespace aq {

template <class T> struct AdjStencil {
  int base[3];
  T fx[3], w[3][3], dw[3][3];
};

template <int DIM, class T>
__device__ __forceinline__ AdjStencil<T> adj_stencil(const T *x, const ParamsT<T> &P) {
  AdjStencil<T> S;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    S.base[a] = 0;
    S.fx[a] = (T)0;
#pragma unroll
    for (int i = 0; i < 3; ++i) S.w[a][i] = S.dw[a][i] = (T)0;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    S.base[a] = stencil_1d<T>(x[a], P, S.fx[a], S.w[a]);
    const T f = S.fx[a];
    S.dw[a][0] = f - (T)1.5;
    S.dw[a][1] = (T)-2 * (f - (T)1);
    S.dw[a][2] = f - (T)0.5;
  }
  return S;
}

// weight of stencil node off and its gradient in fx
template <int DIM, class T>
__device__ __forceinline__ T adj_weight(const AdjStencil<T> &S, const int (&off)[3], T (&g)[3]) {
  T w = (T)1;
#pragma unroll
  for (int a = 0; a < DIM; ++a) w *= S.w[a][off[a]];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g[c] = (T)0;
    if (c < DIM) {
      T v = S.dw[c][off[c]];
#pragma unroll
      for (int a = 0; a < DIM; ++a)
        if (a != c) v *= S.w[a][off[a]];
      g[c] = v;
    }
  }
  return w;
}

template <int DIM>
__device__ __forceinline__ int64_t adj_node(const int (&base)[3], const int (&off)[3], int nb) {
  constexpr int LG = BlockShape<DIM>::LOG, E = BlockShape<DIM>::E;
  const int i = base[0] + off[0], j = base[1] + off[1], k = base[2] + off[2];
  if constexpr (DIM == 3) {
    const int64_t blk = ((int64_t)(i >> LG) * nb + (j >> LG)) * nb + (k >> LG);
    return blk * 64 + (((i & (E - 1)) << 4) | ((j & (E - 1)) << 2) | (k & (E - 1)));
  } else {
    const int64_t blk = (int64_t)(i >> LG) * nb + (j >> LG);
    return blk * 64 + (((i & (E - 1)) << 3) | (j & (E - 1)));
  }
}

// G2P adjoint.  st, adj_next, adj_out: value planes [NC][cap]; grid_v: velocities.
template <int DIM, int MAT, class T>
__global__ void __launch_bounds__(256) k_adj_g2p(const T *__restrict__ st, const T *__restrict__ adj_next,
                                                 T *__restrict__ adj_out, int64_t cap, int64_t n, ParamsT<T> P,
                                                 const vec4_t<T> *__restrict__ grid_v, T *__restrict__ gbar) {
  using CI = Comps<DIM, MAT>;
  constexpr int NC = CI::NC, NJF = NC - CI::CJ, NK = DIM == 3 ? 3 : 1;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  T x[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int a = 0; a < DIM; ++a) x[a] = st[(int64_t)(CI::CX + a) * cap + p];
  const AdjStencil<T> S = adj_stencil<DIM, T>(x, P);
  T nC[DIM * DIM];
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) nC[i] = (T)0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        const int off[3] = {i, j, k};
        T gw[3];
        const T w = adj_weight<DIM, T>(S, off, gw);
        const vec4_t<T> nd = grid_v[adj_node<DIM>(S.base, off, P.nb)];
        const T g[3] = {nd.x, nd.y, nd.z};
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
          for (int b = 0; b < DIM; ++b) nC[a * DIM + b] += P.c4 * w * g[a] * ((T)off[b] - S.fx[b]);
      }
  T xb[3] = {(T)0, (T)0, (T)0}, vb[3] = {(T)0, (T)0, (T)0}, Cb[DIM * DIM], jf[NJF], jfb[NJF];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    xb[a] = adj_next[(int64_t)(CI::CX + a) * cap + p];
    vb[a] = adj_next[(int64_t)(CI::CV + a) * cap + p] + P.dt * xb[a];
  }
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) Cb[i] = adj_next[(int64_t)(CI::CC + i) * cap + p];
#pragma unroll
  for (int i = 0; i < NJF; ++i) jf[i] = st[(int64_t)(CI::CJ + i) * cap + p];
  if constexpr (MAT == AQ_FLUID) {
    T tr = (T)0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) tr += nC[a * DIM + a];
    const T Jb = adj_next[(int64_t)CI::CJ * cap + p];
    jfb[0] = Jb * ((T)1 + P.dt * tr);
#pragma unroll
    for (int a = 0; a < DIM; ++a) Cb[a * DIM + a] += Jb * P.dt * jf[0];
  } else {
    T Fbn[DIM * DIM], M[DIM * DIM];
#pragma unroll
    for (int i = 0; i < DIM * DIM; ++i) Fbn[i] = adj_next[(int64_t)(CI::CJ + i) * cap + p];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) M[a * DIM + b] = (a == b ? (T)1 : (T)0) + P.dt * nC[a * DIM + b];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) {
        T acc = (T)0, acc2 = (T)0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
          acc += M[k * DIM + a] * Fbn[k * DIM + b];
          acc2 += Fbn[a * DIM + k] * jf[b * DIM + k];
        }
        jfb[a * DIM + b] = acc;
        Cb[a * DIM + b] += P.dt * acc2;
      }
  }
  T fxb[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        const int off[3] = {i, j, k};
        T gw[3];
        const T w = adj_weight<DIM, T>(S, off, gw);
        const int64_t node = adj_node<DIM>(S.base, off, P.nb);
        const vec4_t<T> nd = grid_v[node];
        const T g[3] = {nd.x, nd.y, nd.z};
        T u[3] = {(T)0, (T)0, (T)0}, gu = (T)0;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          u[a] = vb[a];
#pragma unroll
          for (int b = 0; b < DIM; ++b) u[a] += P.c4 * Cb[a * DIM + b] * ((T)off[b] - S.fx[b]);
          atomic_add<T>(gbar + node * 4 + a, w * u[a]);
          gu += g[a] * u[a];
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          T t = (T)0;
#pragma unroll
          for (int a = 0; a < DIM; ++a) t += Cb[a * DIM + c] * g[a];
          fxb[c] += gw[c] * gu - P.c4 * w * t;
        }
      }
#pragma unroll
  for (int a = 0; a < DIM; ++a) adj_out[(int64_t)(CI::CX + a) * cap + p] = xb[a] + P.inv_dx * fxb[a];
#pragma unroll
  for (int i = 0; i < NJF; ++i) adj_out[(int64_t)(CI::CJ + i) * cap + p] = jfb[i];
}

// Grid-update adjoint on the active blocks: gbar (adjoint of the velocity)
// becomes the adjoint of (m v, m) in place; pre holds the P2G result.  With
// the collider (3D), first its adjoint: where the node is inside and the
// velocity after the walls moves into the collider (v_rel . n < 0), the
// update v - (v_rel . n) n has Jacobian I - n n^T.
template <int DIM, class T>
__global__ void __launch_bounds__(256) k_adj_grid(ParamsT<T> P, ColliderK K, int32_t nblocks,
                                                  const uint32_t *__restrict__ act,
                                                  const vec4_t<T> *__restrict__ pre, T *__restrict__ gbar) {
  const int32_t b = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (b >= nblocks || !act[b]) return;
  constexpr int E = BlockShape<DIM>::E;
  int bc[3], I[3];
  if constexpr (DIM == 3) {
    bc[0] = b / (P.nb * P.nb);
    bc[1] = (b / P.nb) % P.nb;
    bc[2] = b % P.nb;
    I[0] = bc[0] * E + (lane >> 4);
    I[1] = bc[1] * E + ((lane >> 2) & 3);
    I[2] = bc[2] * E + (lane & 3);
  } else {
    bc[0] = b / P.nb;
    bc[1] = b % P.nb;
    I[0] = bc[0] * E + (lane >> 3);
    I[1] = bc[1] * E + (lane & 7);
    I[2] = 0;
  }
  const int64_t node = (int64_t)b * 64 + lane;
  const vec4_t<T> nd = pre[node];
  const T mv[3] = {nd.x, nd.y, nd.z};
  const T inv_m = (T)1 / (nd.w + P.eps_m);
  T v[3] = {inv_m * mv[0], inv_m * mv[1], inv_m * mv[2]};
  v[1] = v[1] - P.g_dt;
  T gb[3] = {gbar[node * 4 + 0], gbar[node * 4 + 1], gbar[node * 4 + 2]};
  if constexpr (DIM == 3) {
    T n[3], cv[3];
    if (K.on && collider_probe<T>(P, K, I, n, cv)) {
      T vw[3] = {v[0], v[1], v[2]};  // the velocity after the walls, as the forward collides it
      if (P.boundary == AQ_STICKY) {
        bool wall = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) wall = wall || I[d] < P.lb[d] || I[d] > P.ub[d];
        if (wall) vw[0] = vw[1] = vw[2] = (T)0;
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (I[d] < P.lb[d] && vw[d] < (T)0) vw[d] = (T)0;
          if (I[d] > P.ub[d] && vw[d] > (T)0) vw[d] = (T)0;
        }
      }
      const T dn = (vw[0] - cv[0]) * n[0] + (vw[1] - cv[1]) * n[1] + (vw[2] - cv[2]) * n[2];
      if (dn < (T)0) {
        const T ng = gb[0] * n[0] + gb[1] * n[1] + gb[2] * n[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) gb[a] -= ng * n[a];
      }
    }
  }
  if (P.boundary == AQ_STICKY) {
    bool wall = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) wall = wall || I[d] < P.lb[d] || I[d] > P.ub[d];
    if (wall) gb[0] = gb[1] = gb[2] = (T)0;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      if (I[d] < P.lb[d] && v[d] < (T)0) gb[d] = (T)0;
      if (I[d] > P.ub[d] && v[d] > (T)0) gb[d] = (T)0;
    }
  }
  T mb = (T)0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    mb -= gb[a] * mv[a] * inv_m * inv_m;
    gbar[node * 4 + a] = gb[a] * inv_m;
  }
  gbar[node * 4 + 3] = mb;
}

// 2D fixed-corotated stress adjoint: sigma_bar -> F_bar (added).
template <class T>
__device__ __forceinline__ void adj_corotated(const ParamsT<T> &P, const T (&F)[4], const T (&sb)[4], T (&Fb)[4]) {
  const T J = F[0] * F[3] - F[1] * F[2];
  const T x = F[0] + F[3], y = F[2] - F[1];
  const T r = tsqrt<T>(x * x + y * y);
  const T c = x / r, sn = y / r;
  const T R[4] = {c, -sn, sn, c};
  T A[4], Ab[4], Rb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) A[i] = P.two_mu * (F[i] - R[i]);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      Ab[i * 2 + k] = sb[i * 2 + 0] * F[0 * 2 + k] + sb[i * 2 + 1] * F[1 * 2 + k];
      Fb[i * 2 + k] += sb[0 * 2 + i] * A[0 * 2 + k] + sb[1 * 2 + i] * A[1 * 2 + k];
    }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    Fb[i] += P.two_mu * Ab[i];
    Rb[i] = -P.two_mu * Ab[i];
  }
  const T cb = Rb[0] + Rb[3], sbb = Rb[2] - Rb[1];
  const T r3 = r * r * r;
  const T xb = cb * (y * y / r3) - sbb * (x * y / r3);
  const T yb = -cb * (x * y / r3) + sbb * (x * x / r3);
  Fb[0] += xb;
  Fb[3] += xb;
  Fb[2] += yb;
  Fb[1] -= yb;
  const T Jb = P.la_0 * ((T)2 * J - (T)1) * (sb[0] + sb[3]);
  Fb[0] += Jb * F[3];
  Fb[3] += Jb * F[0];
  Fb[1] -= Jb * F[2];
  Fb[2] -= Jb * F[1];
}

// 3D elastic stress adjoint (sigma = mu_0 (F F^T - I) + la_0 log(J) on
// every entry): F_bar += mu_0 (sigma_bar + sigma_bar^T) F
//                      + la_0 (sum_ij sigma_bar_ij) F^-T,  F^-T = cof(F) / J.
template <class T>
__device__ __forceinline__ void adj_elastic3(const ParamsT<T> &P, const T (&F)[9], const T (&sb)[9], T (&Fb)[9]) {
  const T mu = P.two_mu * (T)0.5;
  T ssum = (T)0;
#pragma unroll
  for (int i = 0; i < 9; ++i) ssum += sb[i];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T acc = (T)0;
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += (sb[a * 3 + j] + sb[j * 3 + a]) * F[j * 3 + b];
      Fb[a * 3 + b] += mu * acc;
    }
  // cofactors: cof_ab = (-1)^(a+b) minor_ab; F^-T = cof / det
  const T c00 = F[4] * F[8] - F[5] * F[7], c01 = F[5] * F[6] - F[3] * F[8], c02 = F[3] * F[7] - F[4] * F[6];
  const T c10 = F[2] * F[7] - F[1] * F[8], c11 = F[0] * F[8] - F[2] * F[6], c12 = F[1] * F[6] - F[0] * F[7];
  const T c20 = F[1] * F[5] - F[2] * F[4], c21 = F[2] * F[3] - F[0] * F[5], c22 = F[0] * F[4] - F[1] * F[3];
  const T J = F[0] * c00 + F[1] * c01 + F[2] * c02;
  const T k = P.la_0 * ssum / J;
  const T cof[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
#pragma unroll
  for (int i = 0; i < 9; ++i) Fb[i] += k * cof[i];
}

// P2G adjoint: completes adj_out and adds sum_p adj_out[c][p]^2 to gsum[c].
template <int DIM, int MAT, class T>
__global__ void __launch_bounds__(256) k_adj_p2g(const T *__restrict__ st, T *__restrict__ adj_out, int64_t cap,
                                                 int64_t n, ParamsT<T> P, const T *__restrict__ gbar,
                                                 double *__restrict__ gsum) {
  using CI = Comps<DIM, MAT>;
  constexpr int NC = CI::NC, NJF = NC - CI::CJ, NK = DIM == 3 ? 3 : 1;
  __shared__ double part[4][NC];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  T out[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = (T)0;
  if (p < n) {
    T v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = st[(int64_t)c * cap + p];
    const AdjStencil<T> S = adj_stencil<DIM, T>(v, P);
    T sig[DIM * DIM];
    {
      T aff0[DIM * DIM];
      affine_of<DIM, MAT, T>(v, P, aff0);  // stress_k sigma + m C
#pragma unroll
      for (int i = 0; i < DIM * DIM; ++i) sig[i] = aff0[i];
    }
    const T *aff = sig;
    T vb[3] = {(T)0, (T)0, (T)0}, affb[DIM * DIM], fxb[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int i = 0; i < DIM * DIM; ++i) affb[i] = (T)0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const int off[3] = {i, j, k};
          T gw[3];
          const T w = adj_weight<DIM, T>(S, off, gw);
          const int64_t node = adj_node<DIM>(S.base, off, P.nb);
          T dp[3] = {(T)0, (T)0, (T)0}, mvb[3] = {(T)0, (T)0, (T)0};
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            dp[a] = ((T)off[a] - S.fx[a]) * P.dx;
            mvb[a] = gbar[node * 4 + a];
          }
          const T mb = gbar[node * 4 + 3];
          T contrib = mb * P.p_mass;
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            T q = P.p_mass * v[CI::CV + a];
#pragma unroll
            for (int b = 0; b < DIM; ++b) q += aff[a * DIM + b] * dp[b];
            contrib += mvb[a] * q;
            vb[a] += P.p_mass * w * mvb[a];
#pragma unroll
            for (int b = 0; b < DIM; ++b) affb[a * DIM + b] += w * mvb[a] * dp[b];
          }
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            T t = (T)0;
#pragma unroll
            for (int a = 0; a < DIM; ++a) t += aff[a * DIM + c] * mvb[a];
            fxb[c] += gw[c] * contrib - P.dx * w * t;
          }
        }
    // adjoints from G2P (x and J/F), then the P2G terms
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      out[CI::CX + a] = adj_out[(int64_t)(CI::CX + a) * cap + p] + P.inv_dx * fxb[a];
      out[CI::CV + a] = vb[a];
    }
#pragma unroll
    for (int i = 0; i < DIM * DIM; ++i) out[CI::CC + i] = P.p_mass * affb[i];
    T jfb[NJF];
#pragma unroll
    for (int i = 0; i < NJF; ++i) jfb[i] = adj_out[(int64_t)(CI::CJ + i) * cap + p];
    if constexpr (MAT == AQ_FLUID) {
      T tr = (T)0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) tr += affb[a * DIM + a];
      jfb[0] += P.la * (P.stress_k * tr);
    } else if constexpr (DIM == 3) {
      T F[9], sb[9], Fb[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        F[i] = v[CI::CJ + i];
        sb[i] = P.stress_k * affb[i];
        Fb[i] = (T)0;
      }
      adj_elastic3<T>(P, F, sb, Fb);
#pragma unroll
      for (int i = 0; i < 9; ++i) jfb[i] += Fb[i];
    } else {
      const T F[4] = {v[CI::CJ + 0], v[CI::CJ + 1], v[CI::CJ + 2], v[CI::CJ + 3]};
      T sb[4], Fb[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
      for (int i = 0; i < 4; ++i) sb[i] = P.stress_k * affb[i];
      adj_corotated<T>(P, F, sb, Fb);
#pragma unroll
      for (int i = 0; i < 4; ++i) jfb[i] += Fb[i];
    }
#pragma unroll
    for (int i = 0; i < NJF; ++i) out[CI::CJ + i] = jfb[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) adj_out[(int64_t)c * cap + p] = out[c];
  }
  if (gsum) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double q = (double)out[c] * (double)out[c];
      for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o);
      if (lane == 0) part[wave][c] = q;
    }
    __syncthreads();
    if (threadIdx.x < NC) {
      const int c = threadIdx.x;
      atomicAdd(gsum + c, ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c]);
    }
  }
}

}  // namespace aq
