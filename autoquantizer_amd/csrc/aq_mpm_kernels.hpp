// MLS-MPM substep kernels for gfx950.
//
// Physics follows MPMSolver of the reference (mpm_exps/engine/mpm_solver.py):
//   stencil base/weights  :599-606, :676-678
//   stress + affine       :608-613, compute_cauchy :1012-1031,
//                         polar_decompose2d (TS/python/taichi/lang/linalg_impl.py:9-26)
//   P2G scatter           :615-622
//   grid update           :633-665 (separating walls; sticky walls zero the node)
//   G2P gather + update   :682-694, :702 (J), :750-758 (F)
// The stencil sums are evaluated axis by axis (separable weights, offsets
// centred on the middle node, fused multiply-adds): the same sums as the
// reference's node-by-node loops, about a third of the arithmetic, and
// per-particle results within a few float ulps of the reference's operand
// order.  The codec (aq_codec.hpp) keeps the reference's unfused rounding.
//
// Grid: n^DIM nodes in blocks of 64 (4x4x4 in 3D, 8x8 in 2D), block-major,
// one 16-byte (float4) or 32-byte (double4) record per node:
// (m v_x, m v_y, m v_z, m) after P2G, (v_x, v_y, v_z, m) after the grid update.
// A grid block of 64 nodes is one wavefront in the grid kernels.
#pragma once

#include "aq_codec.hpp"

namespace aq {

template <class T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
template <class T> using vec4_t = typename Vec4<T>::type;

template <class T> struct ParamsT {
  T dx, inv_dx, dt, p_mass, stress_k, la, two_mu, la_0, g_dt, c4, half_pm, g_pm, eps_m;
  T xlo[3], xhi[3];                 // bound clip: lb dx, ub dx per axis
  int32_t n, nb, boundary, clip;    // cells per axis, blocks per axis, wall kind, bound clip on
  int32_t lb[3], ub[3];             // walls per axis (cells)
};

template <int DIM> struct BlockShape;
template <> struct BlockShape<3> { static constexpr int E = 4, LOG = 2; };
template <> struct BlockShape<2> { static constexpr int E = 8, LOG = 3; };

template <int DIM, int MAT> struct Comps {
  static constexpr int CX = 0, CV = DIM, CC = 2 * DIM, CJ = 2 * DIM + DIM * DIM;
  static constexpr int NC = 2 * DIM + DIM * DIM + (MAT == AQ_FLUID ? 1 : DIM * DIM);
};

// ---------------------------------------------------------------------------
// state access: AQ_STORE_QUANT (word planes + layout) or plain value planes
// ---------------------------------------------------------------------------
struct StateK {
  void *ptr;
  int64_t cap;          // plane stride
  const int32_t *ids;   // particle ids (permuted by sorts) or null: key of the dithering noise
  int32_t nw;           // quant: record words
  uint32_t g2p_mask;    // quant: words holding x and J/F (what G2P reads)
  int32_t dither_any;   // quant: some member dithers (ids are read only then)
  int32_t pad;
};

// Whether the compute type is double (a type trait that hipRTC sources can use too).
template <class T> struct IsF64 { static constexpr bool value = false; };
template <> struct IsF64<double> { static constexpr bool value = true; };

// Noise key of a particle: its id when ids are tracked, else its slot.
__device__ __forceinline__ uint32_t noise_slot(const StateK &s, int64_t p) {
  return (s.dither_any && s.ids) ? (uint32_t)s.ids[p] : (uint32_t)p;
}

// Per-thread LDS rows of the record codec (odd stride: distinct banks per
// lane); plain storage needs none.
template <int ST, int NW32> struct RowsOf {
  static constexpr int RS = ST == AQ_STORE_QUANT ? NW32 + 1 : 1;
};

// Raw words of a quantized record fetched ahead of their decode.
template <int ST, int NW32> struct Fetched {
  uint64_t w[ST == AQ_STORE_QUANT ? NW32 / 2 : 1];
};

// Global loads of particle p: the record words in `mask` (quantized) or the
// value planes (plain; with G2P_ONLY only x and J/F, the members G2P reads).
template <int DIM, int ST, int NW32, int NC, class T, bool G2P_ONLY = false>
__device__ __forceinline__ void fetch_state(const StateK &s, int64_t p, uint32_t mask, Fetched<ST, NW32> &f,
                                            T (&v)[NC]) {
  constexpr int CJ = 2 * DIM + DIM * DIM;
  if constexpr (ST == AQ_STORE_QUANT) {
    const uint64_t *w = (const uint64_t *)s.ptr;
#ifdef AQ_JIT_LAYOUT
    constexpr int nw = jit::kNW;
#else
    const int nw = s.nw;
#endif
#pragma unroll
    for (int i = 0; i < NW32 / 2; ++i) {
      f.w[i] = 0ull;
      if (i < nw && ((mask >> i) & 1u)) f.w[i] = __builtin_nontemporal_load(w + (int64_t)i * s.cap + p);
    }
  } else {
    const T *vals = (const T *)s.ptr;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = (G2P_ONLY && c >= DIM && c < CJ) ? (T)0 : vals[(int64_t)c * s.cap + p];
  }
}

// Decode of fetched words through the thread's LDS row (quantized; plain
// values are already in v).
template <int DIM, int ST, int NW32, int NC, class T, bool G2P_ONLY = false>
__device__ __forceinline__ void finish_state(const StateK &s, const CodecK &cod, uint32_t mask,
                                             const Fetched<ST, NW32> &f, uint32_t *row, T (&v)[NC]) {
  if constexpr (ST == AQ_STORE_QUANT) {
    constexpr int CJ = 2 * DIM + DIM * DIM;
#ifdef AQ_JIT_LAYOUT
    // compile-time offsets: fields come straight out of the fetched words
    uint32_t r[NW32];
#pragma unroll
    for (int i = 0; i < NW32 / 2; ++i) {
      r[2 * i] = (uint32_t)f.w[i];
      r[2 * i + 1] = (uint32_t)(f.w[i] >> 32);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      v[c] = (G2P_ONLY && c >= DIM && c < CJ) ? (T)0 : (T)decode_comp(r, comp_load(nullptr, c));
    (void)row;
    (void)mask;
    return;
#endif
#pragma unroll
    for (int i = 0; i < NW32 / 2; ++i) {
      if (i < s.nw && ((mask >> i) & 1u)) {
        row[2 * i] = (uint32_t)f.w[i];
        row[2 * i + 1] = (uint32_t)(f.w[i] >> 32);
      }
    }
    ConstU32 *table = comp_table(cod);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (G2P_ONLY && c >= DIM && c < CJ) {
        v[c] = (T)0;
        continue;
      }
      if (!G2P_ONLY && c % COMP_BATCH == 0) table = comp_table(cod);
      if constexpr (IsF64<T>::value)
        v[c] = decode_row_f64(row, comp_load(table, c), cod.cd + 4 * c);  // f64 compute
      else
        v[c] = (T)decode_row(row, comp_load(table, c));
    }
  }
}

// Load particle p (fetch + decode).
template <int DIM, int ST, int NW32, int NC, class T, bool G2P_ONLY = false>
__device__ __forceinline__ void load_state(const StateK &s, const CodecK &cod, int64_t p, uint32_t mask,
                                           uint32_t *row, T (&v)[NC]) {
  Fetched<ST, NW32> f;
  fetch_state<DIM, ST, NW32, NC, T, G2P_ONLY>(s, p, mask, f, v);
  finish_state<DIM, ST, NW32, NC, T, G2P_ONLY>(s, cod, mask, f, row, v);
}

// Store particle p.  Quantized records are encoded in the thread's LDS row;
// v[] is replaced by the values the stored record decodes to (what the next
// P2G must see).  Plain storage stores v[] exactly.
template <int ST, int NW32, int NC, class T>
__device__ __forceinline__ void store_state(const StateK &s, const CodecK &cod, int64_t p, T (&v)[NC],
                                            uint32_t *row) {
#ifndef AQ_JIT_LAYOUT
  if constexpr (ST == AQ_STORE_QUANT && IsF64<T>::value) {
    // f64 compute of quantized records (the specialised kernel is f32 only)
    uint64_t *w = (uint64_t *)s.ptr;
    double dec[NC];
    if (cod.stream) {
      StreamSink sink{w, s.cap, p, (uint32_t)s.nw, 0u, 0ull};
      encode_record_f64<NC>(v, cod, noise_slot(s, p), sink, dec);
      sink.finish();
    } else {
#pragma unroll
      for (int i = 0; i < NW32; ++i) row[i] = 0u;
      RowSink sink{row};
      encode_record_f64<NC>(v, cod, noise_slot(s, p), sink, dec);
#pragma unroll
      for (int i = 0; i < NW32 / 2; ++i)
        if (i < s.nw)
          __builtin_nontemporal_store((uint64_t)row[2 * i] | ((uint64_t)row[2 * i + 1] << 32),
                                      w + (int64_t)i * s.cap + p);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = dec[c];
    return;
  }
#endif
  if constexpr (ST == AQ_STORE_QUANT) {
    float f[NC], dec[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = (float)v[c];
    uint64_t *w = (uint64_t *)s.ptr;
#ifdef AQ_JIT_LAYOUT
    RegSink<NW32 / 2> sink;
#pragma unroll
    for (int i = 0; i < NW32 / 2; ++i) sink.w[i] = 0ull;
    encode_record<NC>(f, cod, noise_slot(s, p), sink, dec);
#pragma unroll
    for (int i = 0; i < jit::kNW; ++i) __builtin_nontemporal_store(sink.w[i], w + (int64_t)i * s.cap + p);
    (void)row;
#else
    if (cod.stream) {
      StreamSink sink{w, s.cap, p, (uint32_t)s.nw, 0u, 0ull};
      encode_record<NC>(f, cod, noise_slot(s, p), sink, dec);
      sink.finish();
    } else {
#pragma unroll
      for (int i = 0; i < NW32; ++i) row[i] = 0u;
      RowSink sink{row};
      encode_record<NC>(f, cod, noise_slot(s, p), sink, dec);
#pragma unroll
      for (int i = 0; i < NW32 / 2; ++i)
        if (i < s.nw)
          __builtin_nontemporal_store((uint64_t)row[2 * i] | ((uint64_t)row[2 * i + 1] << 32),
                                      w + (int64_t)i * s.cap + p);
    }
#endif
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = (T)dec[c];
  } else {
    T *vals = (T *)s.ptr;
#pragma unroll
    for (int c = 0; c < NC; ++c) vals[(int64_t)c * s.cap + p] = v[c];
  }
}

// ---------------------------------------------------------------------------
// stencil
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T tfloor(T x);
template <> __device__ __forceinline__ float tfloor<float>(float x) { return floorf(x); }
template <> __device__ __forceinline__ double tfloor<double>(double x) { return floor(x); }
template <class T> __device__ __forceinline__ T tmax(T a, T b);
template <> __device__ __forceinline__ float tmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double tmax<double>(double a, double b) { return fmax(a, b); }
template <class T> __device__ __forceinline__ T tmin(T a, T b);
template <> __device__ __forceinline__ float tmin<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ double tmin<double>(double a, double b) { return fmin(a, b); }
template <class T> __device__ __forceinline__ T tsqrt(T a);
template <> __device__ __forceinline__ float tsqrt<float>(float a) { return sqrtf(a); }
template <> __device__ __forceinline__ double tsqrt<double>(double a) { return sqrt(a); }
template <class T> __device__ __forceinline__ T tlog(T a);
template <> __device__ __forceinline__ float tlog<float>(float a) { return logf(a); }
template <> __device__ __forceinline__ double tlog<double>(double a) { return log(a); }

// base = floor(x * inv_dx - 0.5) clamped to [0, n-3] (NaN -> 0); fx; weights.
template <class T>
__device__ __forceinline__ int stencil_1d(T xs, const ParamsT<T> &P, T &fx, T (&w)[3]) {
  T t = xs * P.inv_dx;
  t = t - (T)0.5;
  t = tmax<T>(t, (T)0);
  t = tmin<T>(t, (T)(P.n - 3));
  const int b = (int)tfloor<T>(t);
  fx = xs * P.inv_dx - (T)b;
  const T a = (T)1.5 - fx;
  w[0] = (T)0.5 * (a * a);
  const T bb = fx - (T)1;
  w[1] = (T)0.75 - bb * bb;
  const T c = fx - (T)0.5;
  w[2] = (T)0.5 * (c * c);
  return b;
}

// Node offsets of the three stencil nodes per axis in the blocked layout.
template <int DIM>
__device__ __forceinline__ void axis_offsets(int base, int axis, int nb, int32_t (&o)[3]) {
  constexpr int E = BlockShape<DIM>::E, LG = BlockShape<DIM>::LOG;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int c = base + i;
    const int blk = c >> LG, loc = c & (E - 1);
    int32_t bs, ls;
    if constexpr (DIM == 3) {
      bs = axis == 0 ? nb * nb : (axis == 1 ? nb : 1);
      ls = axis == 0 ? 16 : (axis == 1 ? 4 : 1);
    } else {
      bs = axis == 0 ? nb : 1;
      ls = axis == 0 ? 8 : 1;
    }
    o[i] = blk * bs * 64 + loc * ls;
  }
}

template <int DIM>
__device__ __forceinline__ int32_t block_of_base(const int (&base)[3], int nb) {
  constexpr int LG = BlockShape<DIM>::LOG;
  if constexpr (DIM == 3)
    return ((base[0] >> LG) * nb + (base[1] >> LG)) * nb + (base[2] >> LG);
  else
    return (base[0] >> LG) * nb + (base[1] >> LG);
}

// compute_cauchy * stress coefficient + p_mass * C (mpm_solver.py:608-613, 1012-1031)
template <int DIM, int MAT, class T, int NC>
__device__ __forceinline__ void affine_of(const T (&v)[NC], const ParamsT<T> &P,
                                          T (&aff)[DIM * DIM]) {
  using CI = Comps<DIM, MAT>;
  T cauchy[DIM * DIM];
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) cauchy[i] = (T)0;
  if constexpr (MAT == AQ_FLUID) {
    const T sig = P.la * (v[CI::CJ] - (T)1.0);
#pragma unroll
    for (int i = 0; i < DIM; ++i) cauchy[i * DIM + i] = sig;
  } else if constexpr (DIM == 3) {
    // mu_0 (F F^T - I) + la_0 log(J), the scalar added to every entry as
    // the reference's matrix + scalar broadcast does (mpm_solver.py:1027-1029);
    // J by Matrix.determinant's cofactor expansion (TS/python/taichi/lang/matrix.py:625-628)
    const T *F = &v[CI::CJ];
    const T J = F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
                F[6] * (F[1] * F[5] - F[4] * F[2]);
    const T mu = P.two_mu * (T)0.5;
    const T lg = P.la_0 * tlog<T>(J);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T fft = F[i * 3 + 0] * F[j * 3 + 0];
        fft = fft + F[i * 3 + 1] * F[j * 3 + 1];
        fft = fft + F[i * 3 + 2] * F[j * 3 + 2];
        cauchy[i * 3 + j] = mu * (fft - (i == j ? (T)1 : (T)0)) + lg;
      }
  } else {
    // 2D fixed-corotated
    const T F00 = v[CI::CJ + 0], F01 = v[CI::CJ + 1], F10 = v[CI::CJ + 2], F11 = v[CI::CJ + 3];
    const T J = F00 * F11 - F01 * F10;
    const T x = F00 + F11, y = F10 - F01;
    const T sc = (T)1.0 / tsqrt<T>(x * x + y * y);
    const T c = x * sc, sn = y * sc;
    const T A00 = P.two_mu * (F00 - c), A01 = P.two_mu * (F01 - (-sn));
    const T A10 = P.two_mu * (F10 - sn), A11 = P.two_mu * (F11 - c);
    const T P00 = A00 * F00 + A01 * F01, P01 = A00 * F10 + A01 * F11;
    const T P10 = A10 * F00 + A11 * F01, P11 = A10 * F10 + A11 * F11;
    const T dg = (P.la_0 * J) * (J - (T)1);
    const T og = (((T)0 * P.la_0) * J) * (J - (T)1);
    cauchy[0] = P00 + dg;
    cauchy[1] = P01 + og;
    cauchy[2] = P10 + og;
    cauchy[3] = P11 + dg;
  }
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) aff[i] = P.stress_k * cauchy[i] + P.p_mass * v[CI::CC + i];
}

template <class T> __device__ __forceinline__ void atomic_add(T *a, T v) { unsafeAtomicAdd(a, v); }
template <class T> __device__ __forceinline__ T tfma(T a, T b, T c);
template <> __device__ __forceinline__ float tfma<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <> __device__ __forceinline__ double tfma<double>(double a, double b, double c) { return fma(a, b, c); }
template <class T> __device__ __forceinline__ T tabs(T x) { return x < (T)0 ? -x : x; }

// Per-particle stencil of one particle: base cells, weights, fx.
template <int DIM, class T> struct Stencil {
  int base[3];
  T fx[3];
  T w[3][3];
};

template <int DIM, int MAT, class T, int NC>
__device__ __forceinline__ Stencil<DIM, T> stencil_of(const T (&v)[NC], const ParamsT<T> &P) {
  using CI = Comps<DIM, MAT>;
  Stencil<DIM, T> S;
  S.base[1] = S.base[2] = 0;
  S.fx[1] = S.fx[2] = (T)0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) S.base[a] = stencil_1d<T>(v[CI::CX + a], P, S.fx[a], S.w[a]);
  return S;
}

// G2P of one particle (mpm_solver.py:682-702, 750-758); fetch(i, j, k)
// returns the grid velocity of stencil node (i, j, k) as a vec4.
//   v_new = sum_n w_n g_n
//   C_new = 4 inv_dx sum_n w_n g_n (off_n - fx)^T
//         = 4 inv_dx (M - v_new (fx - 1)^T),  M = sum_n w_n g_n (off_n - 1)^T
// with the sums over the separable weights taken innermost axis first.
template <int DIM, int MAT, class T, int NC, class F>
__device__ __forceinline__ void g2p_apply(const T (&v)[NC], const ParamsT<T> &P,
                                          const Stencil<DIM, T> &S, F &&fetch, T (&out)[NC]) {
  using CI = Comps<DIM, MAT>;
  T nv[3] = {(T)0, (T)0, (T)0};
  T M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = (T)0;
  if constexpr (DIM == 3) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T R[3] = {(T)0, (T)0, (T)0}, Ry[3] = {(T)0, (T)0, (T)0}, Rz[3] = {(T)0, (T)0, (T)0};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T Pz[3] = {(T)0, (T)0, (T)0}, Qz[3] = {(T)0, (T)0, (T)0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const vec4_t<T> nd = fetch(i, j, k);
          const T g[3] = {nd.x, nd.y, nd.z};
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            Pz[a] = tfma<T>(S.w[2][k], g[a], Pz[a]);
            if (k == 0) Qz[a] = tfma<T>(-S.w[2][0], g[a], Qz[a]);
            if (k == 2) Qz[a] = tfma<T>(S.w[2][2], g[a], Qz[a]);
          }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          R[a] = tfma<T>(S.w[1][j], Pz[a], R[a]);
          Rz[a] = tfma<T>(S.w[1][j], Qz[a], Rz[a]);
          if (j == 0) Ry[a] = tfma<T>(-S.w[1][0], Pz[a], Ry[a]);
          if (j == 2) Ry[a] = tfma<T>(S.w[1][2], Pz[a], Ry[a]);
        }
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        nv[a] = tfma<T>(S.w[0][i], R[a], nv[a]);
        M[a][1] = tfma<T>(S.w[0][i], Ry[a], M[a][1]);
        M[a][2] = tfma<T>(S.w[0][i], Rz[a], M[a][2]);
        if (i == 0) M[a][0] = tfma<T>(-S.w[0][0], R[a], M[a][0]);
        if (i == 2) M[a][0] = tfma<T>(S.w[0][2], R[a], M[a][0]);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T R[2] = {(T)0, (T)0}, Ry[2] = {(T)0, (T)0};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const vec4_t<T> nd = fetch(i, j, 0);
        const T g[2] = {nd.x, nd.y};
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          R[a] = tfma<T>(S.w[1][j], g[a], R[a]);
          if (j == 0) Ry[a] = tfma<T>(-S.w[1][0], g[a], Ry[a]);
          if (j == 2) Ry[a] = tfma<T>(S.w[1][2], g[a], Ry[a]);
        }
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        nv[a] = tfma<T>(S.w[0][i], R[a], nv[a]);
        M[a][1] = tfma<T>(S.w[0][i], Ry[a], M[a][1]);
        if (i == 0) M[a][0] = tfma<T>(-S.w[0][0], R[a], M[a][0]);
        if (i == 2) M[a][0] = tfma<T>(S.w[0][2], R[a], M[a][0]);
      }
    }
  }
  T nC[DIM * DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a)
#pragma unroll
    for (int b = 0; b < DIM; ++b) nC[a * DIM + b] = P.c4 * tfma<T>(-nv[a], S.fx[b] - (T)1, M[a][b]);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    out[CI::CX + a] = v[CI::CX + a] + P.dt * nv[a];
    if (P.clip) out[CI::CX + a] = tmax<T>(tmin<T>(out[CI::CX + a], P.xhi[a]), P.xlo[a]);
    out[CI::CV + a] = nv[a];
  }
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) out[CI::CC + i] = nC[i];
  if constexpr (MAT == AQ_FLUID) {
    T tr = nC[0];
#pragma unroll
    for (int a = 1; a < DIM; ++a) tr = tr + nC[a * DIM + a];
    out[CI::CJ] = ((T)1.0 + P.dt * tr) * v[CI::CJ];
  } else {
    T M2[DIM * DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) M2[a * DIM + b] = (a == b ? (T)1 : (T)0) + P.dt * nC[a * DIM + b];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) {
        T acc = M2[a * DIM + 0] * v[CI::CJ + 0 * DIM + b];
#pragma unroll
        for (int k = 1; k < DIM; ++k) acc = acc + M2[a * DIM + k] * v[CI::CJ + k * DIM + b];
        out[CI::CJ + a * DIM + b] = acc;
      }
  }
}

// P2G terms of one particle (mpm_solver.py:608-622).  The contribution to
// stencil node n is
//   w_n (m v + aff dpos_n),  dpos_n = (off_n - fx) dx,   and mass w_n m,
// written as w_n (c + A (off_n - 1)) with A = aff dx, c = m v - A (fx - 1).
// Rows of c and A (and pm) may be scaled by powers of two by the caller.
template <int DIM, class T> struct P2GTerms {
  T c[3];
  T A[3][3];
  T pm;
};

template <int DIM, int MAT, class T, int NC>
__device__ __forceinline__ P2GTerms<DIM, T> p2g_terms(const T (&v)[NC], const ParamsT<T> &P,
                                                      const Stencil<DIM, T> &S) {
  using CI = Comps<DIM, MAT>;
  T aff[DIM * DIM];
  affine_of<DIM, MAT, T>(v, P, aff);
  P2GTerms<DIM, T> R;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R.c[a] = (T)0;
#pragma unroll
    for (int b = 0; b < 3; ++b) R.A[a][b] = (T)0;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    T c = P.p_mass * v[CI::CV + a];
#pragma unroll
    for (int b = 0; b < DIM; ++b) {
      R.A[a][b] = aff[a * DIM + b] * P.dx;
      c = tfma<T>(-R.A[a][b], S.fx[b] - (T)1, c);
    }
    R.c[a] = c;
  }
  R.pm = P.p_mass;
  return R;
}

// Upper bound of |contribution| of one particle to any of its stencil nodes,
// per component (momentum x, y, z; mass): max weight * (|c_a| + sum_b
// |A_ab|), with a 2^-10 margin for the rounding of the contributions
// themselves.  NaN/Inf terms give a non-finite bound.
template <int DIM, class T>
__device__ __forceinline__ void p2g_bounds(const P2GTerms<DIM, T> &R, const Stencil<DIM, T> &S, T (&B)[4]) {
  constexpr T margin = (T)1 + (T)1 / (T)1024;
  T W = (T)1;
#pragma unroll
  for (int a = 0; a < DIM; ++a) W *= tmax<T>(tmax<T>(S.w[a][0], S.w[a][1]), S.w[a][2]);
  // a NaN weight must not vanish in tmax
#pragma unroll
  for (int a = 0; a < DIM; ++a) W = (S.fx[a] == S.fx[a]) ? W : S.fx[a];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    B[a] = (T)0;
    if (a < DIM) {
      T acc = tabs<T>(R.c[a]);
#pragma unroll
      for (int b = 0; b < DIM; ++b) acc += tabs<T>(R.A[a][b]);
      B[a] = W * acc * margin;
    }
  }
  B[3] = W * R.pm * margin;
}

// Scatter of one particle: add(i, j, k, mom[3], m) for the 3^DIM stencil
// nodes, mom = w_n (c + A (off_n - 1)), m = w_n pm.
template <int DIM, class T, class F>
__device__ __forceinline__ void p2g_emit(const P2GTerms<DIM, T> &R, const Stencil<DIM, T> &S, F &&add) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T ci[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) ci[a] = i == 0 ? R.c[a] - R.A[a][0] : (i == 2 ? R.c[a] + R.A[a][0] : R.c[a]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const T wij = S.w[0][i] * S.w[1][j];
      T cij[3];
#pragma unroll
      for (int a = 0; a < DIM; ++a) cij[a] = j == 0 ? ci[a] - R.A[a][1] : (j == 2 ? ci[a] + R.A[a][1] : ci[a]);
      if constexpr (DIM == 3) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const T w = wij * S.w[2][k];
          T mom[3];
#pragma unroll
          for (int a = 0; a < 3; ++a)
            mom[a] = w * (k == 0 ? cij[a] - R.A[a][2] : (k == 2 ? cij[a] + R.A[a][2] : cij[a]));
          add(i, j, k, mom, w * R.pm);
        }
      } else {
        T mom[3] = {wij * cij[0], wij * cij[1], (T)0};
        add(i, j, 0, mom, wij * R.pm);
      }
    }
  }
}

// Where an untiled scatter reports its blocks: the per-block flags and,
// when `list` is set, a list that gets each block once when its flag goes
// from 0 to 1 (`count`: the list's length).
struct FlagSink {
  uint32_t *flags;
  int32_t *list;
  uint32_t *count;
};

// Flag every grid block that the stencil of `base` touches (the base block
// and, per axis, the next one when the stencil crosses a block edge).
template <int DIM>
__device__ __forceinline__ void mark_stencil_blocks(const int (&base)[3], int nb, const FlagSink &fs) {
  constexpr int LG = BlockShape<DIM>::LOG;
  int lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = a < DIM ? base[a] >> LG : 0;
    hi[a] = a < DIM ? (base[a] + 2) >> LG : 0;
  }
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) {
        const int32_t b = DIM == 3 ? (i * nb + j) * nb + k : i * nb + j;
        if (fs.list) {
          if (atomicExch(&fs.flags[b], 1u) == 0u) fs.list[atomicAdd(fs.count, 1u)] = b;
        } else {
          fs.flags[b] = 1u;
        }
      }
}

// P2G of one particle with global atomics.
template <int DIM, int MAT, class T, int NC>
__device__ __forceinline__ void p2g_particle(const T (&v)[NC], const ParamsT<T> &P,
                                             vec4_t<T> *__restrict__ grid, const FlagSink &flags) {
  const Stencil<DIM, T> S = stencil_of<DIM, MAT, T>(v, P);
  int32_t o0[3], o1[3], o2[3] = {0, 0, 0};
  axis_offsets<DIM>(S.base[0], 0, P.nb, o0);
  axis_offsets<DIM>(S.base[1], 1, P.nb, o1);
  if constexpr (DIM == 3) axis_offsets<DIM>(S.base[2], 2, P.nb, o2);
  p2g_emit<DIM, T>(p2g_terms<DIM, MAT, T>(v, P, S), S, [&](int i, int j, int k, const T (&mom)[3], T m) {
    T *node = (T *)(grid + (o0[i] + o1[j] + o2[k]));
#pragma unroll
    for (int a = 0; a < DIM; ++a) atomic_add<T>(node + a, mom[a]);
    atomic_add<T>(node + 3, m);
  });
  mark_stencil_blocks<DIM>(S.base, P.nb, flags);
}

// G2P of one particle reading the grid from global memory.
template <int DIM, int MAT, class T, int NC>
__device__ __forceinline__ void g2p_particle(const T (&v)[NC], const ParamsT<T> &P,
                                             const vec4_t<T> *__restrict__ grid, T (&out)[NC]) {
  const Stencil<DIM, T> S = stencil_of<DIM, MAT, T>(v, P);
  int32_t o0[3], o1[3], o2[3] = {0, 0, 0};
  axis_offsets<DIM>(S.base[0], 0, P.nb, o0);
  axis_offsets<DIM>(S.base[1], 1, P.nb, o1);
  if constexpr (DIM == 3) axis_offsets<DIM>(S.base[2], 2, P.nb, o2);
  g2p_apply<DIM, MAT, T>(v, P, S, [&](int i, int j, int k) { return grid[o0[i] + o1[j] + o2[k]]; }, out);
}

// Thread-block index remapped so each XCD (blocks are dealt round-robin to
// the 8 XCDs) gets a contiguous range of particles: spatially sorted
// particles then hit grid nodes cached in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
  const uint32_t per = nblocks / 8u;
  if (bid >= per * 8u) return bid;  // tail keeps its id
  return (bid & 7u) * per + (bid >> 3);
}

}  // namespace aq
