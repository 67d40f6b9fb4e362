// libaqmpm: HIP kernels and the C ABI declared in include/aq_mpm.h.
// Built for gfx950 only (see autoquantizer_amd/build.py).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "aq_mpm.h"
#include "aq_mpm_kernels.hpp"

using namespace aq;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(AQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

extern "C" const char *aq_last_error(void) { return g_err.c_str(); }
extern "C" const char *aq_version(void) { return "aqmpm 0.1 gfx950"; }

// ---------------------------------------------------------------------------
// layouts
// ---------------------------------------------------------------------------
static int build_layout(const aq_component *comps, int nc, int nw, LayoutK *out) {
  if (nc < 1 || nc > AQ_MAX_COMPS) return fail(AQ_ERR_INVALID, "component count out of range");
  if (nw < 1 || nw > AQ_MAX_WORDS) return fail(AQ_ERR_INVALID, "record words out of range");
  std::memset(out, 0, sizeof(*out));
  out->nc = nc;
  out->nw = nw;
  for (int i = 0; i < nc; ++i) {
    const aq_component &c = comps[i];
    CompK &k = out->c[i];
    if (c.kind < AQ_FIXED || c.kind > AQ_F32) return fail(AQ_ERR_INVALID, "bad component kind");
    const int nb = c.kind == AQ_F32 ? 32 : c.num_bits;
    if (nb < 1 || nb > 32) return fail(AQ_ERR_INVALID, "component bits must be in [1, 32]");
    if (c.bit_offset < 0 || c.bit_offset + nb > nw * 64)
      return fail(AQ_ERR_INVALID, "component outside the record");
    k.kind = (uint8_t)c.kind;
    k.nbits = (uint8_t)nb;
    k.is_signed = c.is_signed ? 1 : 0;
    k.dither = (c.kind == AQ_FIXED && c.use_dithering) ? 1 : 0;
    k.off = (uint16_t)c.bit_offset;
    if (c.kind == AQ_FIXED) {
      if (!(c.scale != 0.0) || !std::isfinite(c.scale) || !std::isfinite(c.offset))
        return fail(AQ_ERR_INVALID, "fixed-point scale must be finite and non-zero");
      k.enc_s = (float)(1.0 / c.scale);
      k.dec_s = (float)c.scale;
      k.offs = (float)c.offset;
    }
    if (c.kind == AQ_FLOAT) {
      if (c.exp_bits < 1 || c.exp_bits > 8) return fail(AQ_ERR_INVALID, "exponent bits must be in [1, 8]");
      if (nb - (c.is_signed ? 1 : 0) > 23) return fail(AQ_ERR_INVALID, "custom float digits exceed 23 bits");
      if (c.exp_offset < 0 || c.exp_offset + c.exp_bits > nw * 64)
        return fail(AQ_ERR_INVALID, "exponent outside the record");
      k.exp_bits = (uint8_t)c.exp_bits;
      k.eoff = (uint16_t)c.exp_offset;
      k.shared = c.shared ? 1 : 0;
      k.conv = (uint32_t)(127 - (1 << (c.exp_bits - 1)) + 1);
    }
  }
  // shared groups: contiguous runs of members with one exponent offset
  for (int i = 0; i < nc; ++i) {
    CompK &k = out->c[i];
    if (!(k.kind == K_FLOAT && k.shared)) continue;
    const bool same_prev = i > 0 && out->c[i - 1].kind == K_FLOAT && out->c[i - 1].shared &&
                           out->c[i - 1].eoff == k.eoff;
    const bool same_next = i + 1 < nc && out->c[i + 1].kind == K_FLOAT && out->c[i + 1].shared &&
                           out->c[i + 1].eoff == k.eoff;
    k.gfirst = same_prev ? 0 : 1;
    k.glast = same_next ? 0 : 1;
    if (same_prev && out->c[i - 1].exp_bits != k.exp_bits)
      return fail(AQ_ERR_INVALID, "shared-exponent members differ in exponent width");
    if (k.gfirst)
      for (int j = 0; j < i; ++j)
        if (out->c[j].kind == K_FLOAT && out->c[j].shared && out->c[j].eoff == k.eoff)
          return fail(AQ_ERR_INVALID, "shared-exponent group members must be consecutive");
  }
  return AQ_OK;
}

struct KeysK {
  uint32_t k[AQ_MAX_COMPS];
};

// Dithering keys of one (seed, step): key_c = h(c ^ h(step ^ h(seed_lo ^ h(seed_hi)))).
static uint32_t host_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
static KeysK make_keys(uint64_t seed, uint32_t step) {
  KeysK K;
  uint32_t k = host_hash32((uint32_t)seed ^ host_hash32((uint32_t)(seed >> 32)));
  k = host_hash32(step ^ k);
  for (int c = 0; c < AQ_MAX_COMPS; ++c) K.k[c] = host_hash32((uint32_t)c ^ k);
  return K;
}

static int nw32_bucket(int nw) { return nw <= 2 ? 4 : nw <= 4 ? 8 : nw <= 8 ? 16 : 32; }

// ---------------------------------------------------------------------------
// codec kernels
// ---------------------------------------------------------------------------
template <int NW32>
__global__ void __launch_bounds__(256) k_pack(LayoutK lay, KeysK keys, const float *__restrict__ vals,
                                              const float *__restrict__ noise, uint64_t *__restrict__ words,
                                              int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t r[NW32];
#pragma unroll
  for (int i = 0; i < NW32; ++i) r[i] = 0u;
  uint32_t gmax = 0u;
  // generic component count: shared-group maxima need the whole group first
  for (int c = 0; c < lay.nc; ++c) {
    const CompK &k = lay.c[c];
    const float v = vals[(int64_t)c * n + p];
    if (k.kind == K_FIXED) {
      float u = 0.0f;
      if (k.dither) u = noise ? noise[(int64_t)c * n + p] : noise_u((uint32_t)p, keys.k[c]);
      put_field(r, k.off, k.nbits, enc_fixed(v, k, u));
    } else if (k.kind == K_FLOAT) {
      if (k.shared) {
        if (k.gfirst) {
          gmax = 0u;
          for (int j = c; j < lay.nc; ++j) {
            const uint32_t e = f32_exp(vals[(int64_t)j * n + p]);
            gmax = e > gmax ? e : gmax;
            if (lay.c[j].glast) break;
          }
          put_field(r, k.eoff, k.exp_bits, shared_exp_field(gmax, k));
        }
        put_field(r, k.off, k.nbits, enc_shared(v, gmax, k));
      } else {
        uint32_t e;
        const uint32_t d = enc_float(v, k, e);
        put_field(r, k.eoff, k.exp_bits, e);
        put_field(r, k.off, k.nbits, d);
      }
    } else if (k.kind == K_F32) {
      put_field(r, k.off, 32u, __float_as_uint(v));
    } else {
      put_field(r, k.off, k.nbits, cvt_sat(v, k.is_signed));
    }
  }
  for (int i = 0; i < lay.nw; ++i)
    words[(int64_t)i * n + p] = (uint64_t)r[2 * i] | ((uint64_t)r[2 * i + 1] << 32);
}

template <int NW32>
__global__ void __launch_bounds__(256) k_unpack(LayoutK lay, const uint64_t *__restrict__ words,
                                                float *__restrict__ vals, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t r[NW32];
#pragma unroll
  for (int i = 0; i < NW32; ++i) r[i] = 0u;
  for (int i = 0; i < lay.nw; ++i) {
    const uint64_t x = words[(int64_t)i * n + p];
    r[2 * i] = (uint32_t)x;
    r[2 * i + 1] = (uint32_t)(x >> 32);
  }
  for (int c = 0; c < lay.nc; ++c) vals[(int64_t)c * n + p] = decode_comp(r, lay.c[c]);
}

// Masked store of the components selected by `mask` (bit c = component c)
// into existing records: vals[j * count + t] is the value of the j-th
// selected component for record idx[t].  Unselected members of a
// shared-exponent group that has a selected member are reconstructed from
// the record and re-encoded with the group's new exponent.
template <int NW32>
__global__ void __launch_bounds__(256) k_store_members(LayoutK lay, KeysK keys, uint32_t mask,
                                                       uint64_t *__restrict__ words, int64_t n,
                                                       const int64_t *__restrict__ idx,
                                                       const float *__restrict__ vals, int64_t count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t p = idx[t];
  if (p < 0 || p >= n) return;
  uint32_t r[NW32];
#pragma unroll
  for (int i = 0; i < NW32; ++i) r[i] = 0u;
  for (int i = 0; i < lay.nw; ++i) {
    const uint64_t x = words[(int64_t)i * n + p];
    r[2 * i] = (uint32_t)x;
    r[2 * i + 1] = (uint32_t)(x >> 32);
  }
  int j = 0;
  for (int c = 0; c < lay.nc; ++c) {
    if (!((mask >> c) & 1u)) continue;
    const CompK &k = lay.c[c];
    const float v = vals[(int64_t)j * count + t];
    ++j;
    if (k.kind == K_FLOAT && k.shared) continue;  // groups below
    if (k.kind == K_FIXED) {
      const float u = k.dither ? noise_u((uint32_t)p, keys.k[c]) : 0.0f;
      set_field(r, k.off, k.nbits, enc_fixed(v, k, u));
    } else if (k.kind == K_FLOAT) {
      uint32_t e;
      const uint32_t d = enc_float(v, k, e);
      set_field(r, k.eoff, k.exp_bits, e);
      set_field(r, k.off, k.nbits, d);
    } else if (k.kind == K_F32) {
      set_field(r, k.off, 32u, __float_as_uint(v));
    } else {
      set_field(r, k.off, k.nbits, cvt_sat(v, k.is_signed));
    }
  }
  // shared-exponent groups touched by the mask
  for (int g0 = 0; g0 < lay.nc; ++g0) {
    const CompK &k0 = lay.c[g0];
    if (!(k0.kind == K_FLOAT && k0.shared && k0.gfirst)) continue;
    int g1 = g0;
    while (!lay.c[g1].glast) ++g1;
    uint32_t gm = 0u;
    for (int c = g0; c <= g1; ++c) gm |= (mask >> c) & 1u;
    if (!gm) continue;
    float gv[AQ_MAX_COMPS];
    uint32_t emax = 0u;
    for (int c = g0; c <= g1; ++c) {
      float v;
      if ((mask >> c) & 1u) {
        int jj = 0;
        for (int q = 0; q < c; ++q) jj += (mask >> q) & 1u;
        v = vals[(int64_t)jj * count + t];
      } else {
        v = decode_comp(r, lay.c[c]);
      }
      gv[c - g0] = v;
      const uint32_t e = f32_exp(v);
      emax = e > emax ? e : emax;
    }
    set_field(r, k0.eoff, k0.exp_bits, shared_exp_field(emax, k0));
    for (int c = g0; c <= g1; ++c)
      set_field(r, lay.c[c].off, lay.c[c].nbits, enc_shared(gv[c - g0], emax, lay.c[c]));
  }
  for (int i = 0; i < lay.nw; ++i)
    words[(int64_t)i * n + p] = (uint64_t)r[2 * i] | ((uint64_t)r[2 * i + 1] << 32);
}

template <template <int> class K, class... A>
static int launch_codec(int nw, int64_t n, hipStream_t st, A... args) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (blocks == 0) return AQ_OK;
  switch (nw32_bucket(nw)) {
    case 4: hipLaunchKernelGGL(K<4>::fn, dim3(blocks), dim3(256), 0, st, args...); break;
    case 8: hipLaunchKernelGGL(K<8>::fn, dim3(blocks), dim3(256), 0, st, args...); break;
    case 16: hipLaunchKernelGGL(K<16>::fn, dim3(blocks), dim3(256), 0, st, args...); break;
    default: hipLaunchKernelGGL(K<32>::fn, dim3(blocks), dim3(256), 0, st, args...); break;
  }
  LAUNCH_CHECK();
  return AQ_OK;
}
template <int NW32> struct PackK { static constexpr auto fn = k_pack<NW32>; };
template <int NW32> struct UnpackK { static constexpr auto fn = k_unpack<NW32>; };
template <int NW32> struct StoreK { static constexpr auto fn = k_store_members<NW32>; };

extern "C" int aq_pack(const aq_component *comps, int ncomp, int nwords, const float *vals,
                       const float *noise, uint64_t seed, uint32_t step, uint64_t *words,
                       int64_t n, void *stream) {
  LayoutK lay;
  int rc = build_layout(comps, ncomp, nwords, &lay);
  if (rc) return rc;
  if (n < 0 || n > 0xFFFFFFFFll) return fail(AQ_ERR_INVALID, "n out of range");
  if (n && (!vals || !words)) return fail(AQ_ERR_INVALID, "null buffer");
  return launch_codec<PackK>(nwords, n, (hipStream_t)stream, lay, make_keys(seed, step), vals, noise,
                             words, n);
}

extern "C" int aq_unpack(const aq_component *comps, int ncomp, int nwords, const uint64_t *words,
                         float *vals, int64_t n, void *stream) {
  LayoutK lay;
  int rc = build_layout(comps, ncomp, nwords, &lay);
  if (rc) return rc;
  if (n < 0) return fail(AQ_ERR_INVALID, "n out of range");
  if (n && (!vals || !words)) return fail(AQ_ERR_INVALID, "null buffer");
  return launch_codec<UnpackK>(nwords, n, (hipStream_t)stream, lay, words, vals, n);
}

extern "C" int aq_store_members(const aq_component *comps, int ncomp, int nwords, uint64_t *words,
                                int64_t n, uint32_t mask, const int64_t *idx, const float *vals,
                                int64_t count, uint64_t seed, uint32_t step, void *stream) {
  LayoutK lay;
  int rc = build_layout(comps, ncomp, nwords, &lay);
  if (rc) return rc;
  if (ncomp < 32 && (mask >> ncomp)) return fail(AQ_ERR_INVALID, "mask selects a missing component");
  if (count < 0 || n < 0) return fail(AQ_ERR_INVALID, "negative size");
  if (count && (!idx || !vals || !words)) return fail(AQ_ERR_INVALID, "null buffer");
  if (!mask || !count) return AQ_OK;
  return launch_codec<StoreK>(nwords, count, (hipStream_t)stream, lay, make_keys(seed, step), mask, words, n,
                              idx, vals, count);
}

// Placement (see aq_layout_plan in include/aq_mpm.h).
extern "C" int aq_layout_plan(const aq_container *cs, int ncont, const aq_place *pl, int np,
                              aq_component *out, int *nwords) {
  if (!cs || !pl || !out || !nwords || ncont < 1 || np < 1 || np > 4096)
    return fail(AQ_ERR_INVALID, "bad layout arguments");
  int64_t cursor = 0;
  std::memset(out, 0, sizeof(aq_component) * (size_t)np);
  for (int ci = 0; ci < ncont; ++ci) {
    const aq_container &c = cs[ci];
    int64_t align, cap;
    if (c.kind == AQ_BIT_PACK) {
      align = 64;
      cap = -1;
    } else if (c.kind == AQ_BIT_STRUCT) {
      if (c.num_bits != 8 && c.num_bits != 16 && c.num_bits != 32 && c.num_bits != 64)
        return fail(AQ_ERR_INVALID, "bit_struct size must be 8, 16, 32 or 64");
      align = c.num_bits;
      cap = c.num_bits;
    } else {
      return fail(AQ_ERR_INVALID, "bad container kind");
    }
    cursor = (cursor + align - 1) / align * align;
    const int64_t start = cursor;
    int64_t off = 0;
    int last_group = -1, group_size = 0;
    int64_t group_exp = -1;
    int group_first = -1;
    for (int i = 0; i < np; ++i) {
      const aq_place &p = pl[i];
      if (p.container != ci) continue;
      if (p.kind < AQ_FIXED || p.kind > AQ_F32) return fail(AQ_ERR_INVALID, "bad member kind");
      const int nb = p.kind == AQ_F32 ? 32 : p.num_bits;
      if (nb < 1 || nb > 32) return fail(AQ_ERR_INVALID, "member bits must be in [1, 32]");
      aq_component &o = out[i];
      o.kind = p.kind;
      o.num_bits = nb;
      o.is_signed = p.is_signed ? 1 : 0;
      o.use_dithering = p.use_dithering && p.kind == AQ_FIXED;
      o.scale = p.kind == AQ_FIXED ? p.scale : 1.0;
      o.offset = p.kind == AQ_FIXED ? p.offset : 0.0;
      if (p.kind == AQ_FLOAT) {
        if (p.exp_bits < 1 || p.exp_bits > 8) return fail(AQ_ERR_INVALID, "exponent bits must be in [1, 8]");
        const bool in_group = p.shared_group >= 0 && p.shared_group == last_group;
        if (in_group) {
          if (pl[group_first].exp_bits != p.exp_bits)
            return fail(AQ_ERR_INVALID, "members of a shared exponent need one exponent width");
          if (i > 0 && pl[i - 1].shared_group != p.shared_group)
            return fail(AQ_ERR_INVALID, "members of a shared exponent must be consecutive");
          ++group_size;
        } else {
          // exponent placed before its first user (snode_expr_utils.cpp:44-66)
          group_exp = start + off;
          off += p.exp_bits;
          last_group = p.shared_group;
          group_first = i;
          group_size = 1;
        }
        o.exp_bits = p.exp_bits;
        o.exp_offset = (int32_t)group_exp;
        o.shared = p.shared_group >= 0 ? 1 : 0;
      } else {
        last_group = -1;
      }
      o.bit_offset = (int32_t)(start + off);
      off += nb;
      if (cap > 0 && off > cap) return fail(AQ_ERR_INVALID, "bit_struct overflows its container");
    }
    (void)group_size;
    cursor = start + (cap > 0 ? cap : (off + 63) / 64 * 64);
  }
  // a "shared" group of one member is an ordinary custom float
  for (int i = 0; i < np; ++i) {
    if (out[i].kind != AQ_FLOAT || !out[i].shared) continue;
    int members = 0;
    for (int j = 0; j < np; ++j)
      if (out[j].kind == AQ_FLOAT && out[j].shared && out[j].exp_offset == out[i].exp_offset) ++members;
    if (members == 1) out[i].shared = 0;
  }
  const int64_t w = (cursor + 63) / 64;
  if (w > AQ_MAX_WORDS) return fail(AQ_ERR_CAPACITY, "record exceeds AQ_MAX_WORDS words");
  *nwords = (int)w;
  return AQ_OK;
}

// ---------------------------------------------------------------------------
// MPM kernels
// ---------------------------------------------------------------------------
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_p2g(StateK s, LayoutK lay, ParamsT<T> P, int64_t n,
                                             vec4_t<T> *__restrict__ grid, uint32_t *__restrict__ flags) {
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC];
  load_state<ST, NW32, NC, T>(s, lay, p, 0xFFFFFFFFu, v);
  const int32_t blk = p2g_particle<DIM, MAT, T>(v, P, grid);
  flags[blk] = 1u;
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_g2p(StateK s, LayoutK lay, KeysK keys, ParamsT<T> P, int64_t n,
                                             const vec4_t<T> *__restrict__ grid) {
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC], out[NC];
  load_state<ST, NW32, NC, T>(s, lay, p, s.g2p_mask, v);
  g2p_particle<DIM, MAT, T>(v, P, grid, out);
  uint32_t r[NW32 > 0 ? NW32 : 1];
  store_state<ST, NW32, NC, T>(s, lay, keys.k, p, out, r);
}

// G2P of substep t (reading grid_v) fused with P2G of substep t+1
// (scattering into grid_s): the state is read once and written once.
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_g2p2g(StateK s, LayoutK lay, KeysK keys, ParamsT<T> P, int64_t n,
                                               const vec4_t<T> *__restrict__ grid_v,
                                               vec4_t<T> *__restrict__ grid_s, uint32_t *__restrict__ flags_s) {
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC], out[NC];
  load_state<ST, NW32, NC, T>(s, lay, p, s.g2p_mask, v);
  g2p_particle<DIM, MAT, T>(v, P, grid_v, out);
  uint32_t r[NW32 > 0 ? NW32 : 1];
  store_state<ST, NW32, NC, T>(s, lay, keys.k, p, out, r);
  if constexpr (ST == AQ_STORE_QUANT) {
    // the next P2G sees what storage holds: decode the record just written
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = (T)decode_comp(r, lay.c[c]);
  }
  const int32_t blk = p2g_particle<DIM, MAT, T>(out, P, grid_s);
  flags_s[blk] = 1u;
}

// Grid update (mpm_solver.py:633-665) on one 64-node block per wavefront.
// A block is active when a particle based in it or in a lower neighbour
// block scattered this substep.  Optionally zeroes the blocks of another
// grid that were active (the grid consumed by the previous G2P).
template <int DIM, class T>
__global__ void __launch_bounds__(256) k_grid_update(ParamsT<T> P, int32_t nblocks, vec4_t<T> *__restrict__ grid,
                                                     uint32_t *__restrict__ flags, uint32_t *__restrict__ act,
                                                     vec4_t<T> *__restrict__ other, uint32_t *__restrict__ oflags,
                                                     uint32_t *__restrict__ oact) {
  const int32_t b = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (b >= nblocks) return;
  if (other) {
    const uint32_t was = oact[b];
    if (was) {
      vec4_t<T> z;
      z.x = z.y = z.z = z.w = (T)0;
      other[(int64_t)b * 64 + lane] = z;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0 && was) {
      oflags[b] = 0u;
      oact[b] = 0u;
    }
  }
  const int nb = P.nb;
  int bc[3];
  if constexpr (DIM == 3) {
    bc[0] = b / (nb * nb);
    bc[1] = (b / nb) % nb;
    bc[2] = b % nb;
  } else {
    bc[0] = b / nb;
    bc[1] = b % nb;
    bc[2] = 0;
  }
  uint32_t a = 0u;
  constexpr int ND = DIM == 3 ? 8 : 4;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const int di = (d >> 0) & 1, dj = (d >> 1) & 1, dk = DIM == 3 ? ((d >> 2) & 1) : 0;
    const int i = bc[0] - di, j = bc[1] - dj, k = bc[2] - dk;
    if (i < 0 || j < 0 || k < 0) continue;
    const int32_t nbid = DIM == 3 ? (i * nb + j) * nb + k : i * nb + j;
    a |= flags[nbid];
  }
  if (!a) return;
  if (lane == 0) act[b] = 1u;
  int I[3];
  if constexpr (DIM == 3) {
    I[0] = bc[0] * 4 + (lane >> 4);
    I[1] = bc[1] * 4 + ((lane >> 2) & 3);
    I[2] = bc[2] * 4 + (lane & 3);
  } else {
    I[0] = bc[0] * 8 + (lane >> 3);
    I[1] = bc[1] * 8 + (lane & 7);
    I[2] = 0;
  }
  vec4_t<T> nd = grid[(int64_t)b * 64 + lane];
  const T inv_m = (T)1 / (nd.w + P.eps_m);
  T v[3] = {inv_m * nd.x, inv_m * nd.y, inv_m * nd.z};
  v[1] = v[1] - P.g_dt;
  const int lb = P.bound, ub = P.n - P.bound;
  if (P.boundary == AQ_STICKY) {
    bool wall = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) wall = wall || I[d] < lb || I[d] > ub;
    if (wall) v[0] = v[1] = v[2] = (T)0;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      if (I[d] < lb && v[d] < (T)0) v[d] = (T)0;
      if (I[d] > ub && v[d] > (T)0) v[d] = (T)0;
    }
  }
  nd.x = v[0];
  nd.y = v[1];
  nd.z = DIM == 3 ? v[2] : (T)0;
  grid[(int64_t)b * 64 + lane] = nd;
}

template <class T>
__global__ void __launch_bounds__(256) k_grid_clear(int32_t nblocks, vec4_t<T> *__restrict__ grid,
                                                    uint32_t *__restrict__ flags, uint32_t *__restrict__ act) {
  const int32_t b = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (b >= nblocks) return;
  const uint32_t was = act[b];
  if (was) {
    vec4_t<T> z;
    z.x = z.y = z.z = z.w = (T)0;
    grid[(int64_t)b * 64 + lane] = z;
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    flags[b] = 0u;
    act[b] = 0u;
  }
}

__global__ void k_fill_u32(uint32_t *a, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---- sort by grid block of the stencil base --------------------------------
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_sort_keys(StateK s, LayoutK lay, ParamsT<T> P, int64_t n,
                                                   uint32_t *__restrict__ key, uint32_t *__restrict__ rank,
                                                   uint32_t *__restrict__ counts) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int base[3] = {0, 0, 0};
  T fx, w[3];
  if constexpr (ST == AQ_STORE_QUANT) {
    uint32_t r[NW32];
    load_words(s, p, s.g2p_mask, r);
#pragma unroll
    for (int a = 0; a < DIM; ++a) base[a] = stencil_1d<T>((T)decode_comp(r, lay.c[a]), P, fx, w);
  } else {
    const T *vals = (const T *)s.ptr;
#pragma unroll
    for (int a = 0; a < DIM; ++a) base[a] = stencil_1d<T>(vals[(int64_t)a * s.cap + p], P, fx, w);
  }
  const uint32_t k = (uint32_t)block_of_base<DIM>(base, P.nb);
  key[p] = k;
  rank[p] = atomicAdd(counts + k, 1u);
}

// Exclusive scan of counts[0..m) into offsets[0..m] by one workgroup of 1024.
__global__ void __launch_bounds__(1024) k_scan(const uint32_t *__restrict__ counts, uint32_t *__restrict__ offsets,
                                               int32_t m) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int per = (m + 1023) / 1024;
  const int lo = t * per, hi = lo + per < m ? lo + per : m;
  uint32_t sum = 0;
  for (int i = lo; i < hi; ++i) sum += counts[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t x = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int i = lo; i < hi; ++i) {
    offsets[i] = run;
    run += counts[i];
  }
  if (t == 1023) offsets[m] = part[1023];
}

template <class E>
__global__ void __launch_bounds__(256) k_sort_scatter(const E *__restrict__ src, E *__restrict__ dst, int32_t planes,
                                                      int64_t cap, int64_t n, const uint32_t *__restrict__ key,
                                                      const uint32_t *__restrict__ rank,
                                                      const uint32_t *__restrict__ offsets,
                                                      const int32_t *__restrict__ ids_in, int32_t *__restrict__ ids_out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t q = (int64_t)offsets[key[p]] + rank[p];
  for (int i = 0; i < planes; ++i) dst[(int64_t)i * cap + q] = src[(int64_t)i * cap + p];
  if (ids_in) ids_out[q] = ids_in[p];
}

// ---- reductions ----------------------------------------------------------------
__device__ __forceinline__ void atomic_max_f32(float *a, float v) {
  if (v >= 0.0f) atomicMax((int *)a, __float_as_int(v));
  else atomicMin((unsigned int *)a, __float_as_uint(v));
}
__device__ __forceinline__ void atomic_min_f32(float *a, float v) {
  if (v >= 0.0f) atomicMin((int *)a, __float_as_int(v));
  else atomicMax((unsigned int *)a, __float_as_uint(v));
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_limits(StateK s, LayoutK lay, int64_t n, float *__restrict__ mxmn) {
  constexpr int NC = Comps<DIM, MAT>::NC;
  __shared__ float smx[NC][8], smn[NC][8];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  T v[NC];
  float mx[NC], mn[NC];
  if (p < n) {
    load_state<ST, NW32, NC, T>(s, lay, p, 0xFFFFFFFFu, v);
#pragma unroll
    for (int c = 0; c < NC; ++c) mx[c] = mn[c] = (float)v[c];
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      mx[c] = -INFINITY;
      mn[c] = INFINITY;
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    for (int o = 32; o > 0; o >>= 1) {
      mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o));
      mn[c] = fminf(mn[c], __shfl_xor(mn[c], o));
    }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      smx[c][wv] = mx[c];
      smn[c][wv] = mn[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const int c = threadIdx.x;
    float a = -INFINITY, b = INFINITY;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a = fmaxf(a, smx[c][w]);
      b = fminf(b, smn[c][w]);
    }
    atomic_max_f32(mxmn + c, a);
    atomic_min_f32(mxmn + NC + c, b);
  }
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_energy(StateK s, LayoutK lay, ParamsT<T> P, int64_t n, double *__restrict__ out) {
  constexpr int NC = Comps<DIM, MAT>::NC;
  using CI = Comps<DIM, MAT>;
  __shared__ double part[8];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (p < n) {
    T v[NC];
    load_state<ST, NW32, NC, T>(s, lay, p, 0xFFFFFFFFu, v);
    T v2 = (T)0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) v2 = v2 + v[CI::CV + a] * v[CI::CV + a];
    if constexpr (DIM == 3) {
      e = (double)(P.half_pm * v2);
      e += (double)(P.g_pm * v[CI::CX + 1]);
    } else {
      e = (double)v2;
    }
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    atomicAdd(out, t);
  }
}

// ---------------------------------------------------------------------------
// simulation handle
// ---------------------------------------------------------------------------
struct aq_sim {
  aq_sim_desc d;
  LayoutK lay;
  ParamsT<float> pf;
  ParamsT<double> pd;
  int64_t nodes;
  int32_t nblocks;
  int32_t node_bytes;
  void *grid[2];
  uint32_t *flags[2], *act[2];
  bool dirty[2];
  int cur;  // grid holding the last velocities
  uint32_t *skey, *srank, *scount, *soff;
  void *tmp_state;
  int32_t *tmp_ids;
  int64_t ws_bytes;
};

static int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

static int check_desc(const aq_sim_desc *d) {
  if (!d) return fail(AQ_ERR_INVALID, "null descriptor");
  if (d->dim != 2 && d->dim != 3) return fail(AQ_ERR_INVALID, "dim must be 2 or 3");
  if (d->material != AQ_FLUID && d->material != AQ_ELASTIC) return fail(AQ_ERR_INVALID, "bad material");
  if (d->material == AQ_ELASTIC && d->dim != 2)
    return fail(AQ_ERR_INVALID, "fixed-corotated elasticity is implemented in 2D");
  const int e = d->dim == 3 ? 4 : 8;
  if (d->n_grid < 8 || d->n_grid % e) return fail(AQ_ERR_INVALID, "n_grid must be a multiple of the block edge");
  if (d->dim == 3 && d->n_grid > 1024) return fail(AQ_ERR_INVALID, "n_grid too large");
  if (d->storage < AQ_STORE_QUANT || d->storage > AQ_STORE_F64) return fail(AQ_ERR_INVALID, "bad storage");
  const int nc = 2 * d->dim + d->dim * d->dim + (d->material == AQ_FLUID ? 1 : d->dim * d->dim);
  if (d->ncomp != nc) return fail(AQ_ERR_INVALID, "ncomp does not match dim/material");
  if (d->capacity < 0 || d->capacity > 0x7FFFFFFFll) return fail(AQ_ERR_INVALID, "capacity out of range");
  if (d->boundary != AQ_SEPARATING && d->boundary != AQ_STICKY) return fail(AQ_ERR_INVALID, "bad boundary");
  return AQ_OK;
}

static void ws_plan(const aq_sim_desc *d, int64_t *offs /*[10]*/, int64_t *total) {
  const int64_t n = d->n_grid;
  const int64_t nodes = d->dim == 3 ? n * n * n : n * n;
  const int64_t nblocks = nodes / 64;
  const int64_t node_bytes = d->storage == AQ_STORE_F64 ? 32 : 16;
  const int64_t cap = d->capacity;
  const int64_t state_bytes = d->storage == AQ_STORE_QUANT ? cap * 8 * d->nwords
                              : cap * d->ncomp * (d->storage == AQ_STORE_F64 ? 8 : 4);
  int64_t o = 0;
  offs[0] = o; o = align256(o + nodes * node_bytes);  // grid 0
  offs[1] = o; o = align256(o + nodes * node_bytes);  // grid 1
  offs[2] = o; o = align256(o + nblocks * 4 * 4);     // flags0, flags1, act0, act1
  offs[3] = o; o = align256(o + cap * 4);             // sort keys
  offs[4] = o; o = align256(o + cap * 4);             // sort ranks
  offs[5] = o; o = align256(o + (nblocks + 1) * 4);   // counts
  offs[6] = o; o = align256(o + (nblocks + 1) * 4);   // offsets
  offs[7] = o; o = align256(o + state_bytes);         // sort temp state
  offs[8] = o; o = align256(o + cap * 4);             // sort temp ids
  *total = o;
}

extern "C" int64_t aq_sim_workspace_bytes(const aq_sim_desc *d) {
  if (check_desc(d)) return -1;
  int64_t offs[10], total;
  ws_plan(d, offs, &total);
  return total;
}

template <class T>
static ParamsT<T> make_params(const aq_sim_desc *d) {
  ParamsT<T> P;
  P.dx = (T)d->dx;
  P.inv_dx = (T)d->inv_dx;
  P.dt = (T)d->dt;
  P.p_mass = (T)d->p_mass;
  P.stress_k = (T)(-(d->dt * d->p_vol * 4 * d->inv_dx * d->inv_dx));
  P.la = (T)d->la;
  P.two_mu = (T)(2 * d->mu_0);
  P.la_0 = (T)d->la_0;
  P.g_dt = (T)(d->dt * d->gravity);
  P.c4 = (T)(4 * d->inv_dx);
  P.half_pm = (T)(0.5 * d->p_mass);
  P.g_pm = (T)(d->gravity * d->p_mass);
  P.eps_m = (T)1e-10;
  P.n = d->n_grid;
  P.nb = d->n_grid / (d->dim == 3 ? 4 : 8);
  P.bound = d->bound;
  P.boundary = d->boundary;
  return P;
}

extern "C" int aq_sim_create(const aq_sim_desc *d, void *ws, int64_t ws_bytes, aq_sim **out) {
  int rc = check_desc(d);
  if (rc) return rc;
  if (!out) return fail(AQ_ERR_INVALID, "null output");
  int64_t offs[10], total;
  ws_plan(d, offs, &total);
  if (!ws || ws_bytes < total) return fail(AQ_ERR_CAPACITY, "workspace too small");
  aq_sim *s = new aq_sim();
  s->d = *d;
  if (d->storage == AQ_STORE_QUANT) {
    rc = build_layout(d->comps, d->ncomp, d->nwords, &s->lay);
    if (rc) {
      delete s;
      return rc;
    }
  } else {
    std::memset(&s->lay, 0, sizeof(s->lay));
  }
  s->pf = make_params<float>(d);
  s->pd = make_params<double>(d);
  const int64_t n = d->n_grid;
  s->nodes = d->dim == 3 ? n * n * n : n * n;
  s->nblocks = (int32_t)(s->nodes / 64);
  s->node_bytes = d->storage == AQ_STORE_F64 ? 32 : 16;
  char *base = (char *)ws;
  s->grid[0] = base + offs[0];
  s->grid[1] = base + offs[1];
  uint32_t *fl = (uint32_t *)(base + offs[2]);
  s->flags[0] = fl;
  s->flags[1] = fl + s->nblocks;
  s->act[0] = fl + 2 * s->nblocks;
  s->act[1] = fl + 3 * s->nblocks;
  s->skey = (uint32_t *)(base + offs[3]);
  s->srank = (uint32_t *)(base + offs[4]);
  s->scount = (uint32_t *)(base + offs[5]);
  s->soff = (uint32_t *)(base + offs[6]);
  s->tmp_state = base + offs[7];
  s->tmp_ids = (int32_t *)(base + offs[8]);
  s->ws_bytes = ws_bytes;
  s->dirty[0] = s->dirty[1] = false;
  s->cur = 0;
  // grids, flags and activity start at zero
  hipError_t e = hipMemset(ws, 0, (size_t)offs[3]);
  if (e != hipSuccess) {
    delete s;
    return fail(AQ_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  *out = s;
  return AQ_OK;
}

extern "C" int aq_sim_destroy(aq_sim *s) {
  delete s;
  return AQ_OK;
}

static StateK make_state(const aq_sim *s, const void *state) {
  StateK k;
  k.ptr = (void *)state;
  k.cap = s->d.capacity;
  k.nw = s->d.nwords;
  uint32_t m = 0u;
  if (s->d.storage == AQ_STORE_QUANT) {
    const int d = s->d.dim;
    const int cj = 2 * d + d * d;
    auto mark = [&](int c) {
      const CompK &cc = s->lay.c[c];
      const int nb = cc.kind == K_F32 ? 32 : cc.nbits;
      for (int b = cc.off; b < cc.off + nb; b += 1) m |= 1u << (b >> 6);
      if (cc.kind == K_FLOAT)
        for (int b = cc.eoff; b < cc.eoff + cc.exp_bits; b += 1) m |= 1u << (b >> 6);
    };
    for (int c = 0; c < d; ++c) mark(c);
    for (int c = cj; c < s->d.ncomp; ++c) mark(c);
  }
  k.g2p_mask = m;
  return k;
}

// Dispatch over (dim, material, storage, record bucket): calls f(Tag<...>{}).
template <int D_, int M_, int S_, int W_, class T_> struct Tag {
  static constexpr int DIM = D_, MAT = M_, ST = S_, NW32 = W_;
  using T = T_;
};

template <int D_, int M_, class F>
static void dispatch_quant(int bucket, F &&f) {
  if (bucket == 4) f(Tag<D_, M_, AQ_STORE_QUANT, 4, float>{});
  else if (bucket == 8) f(Tag<D_, M_, AQ_STORE_QUANT, 8, float>{});
  else if (bucket == 16) f(Tag<D_, M_, AQ_STORE_QUANT, 16, float>{});
  else f(Tag<D_, M_, AQ_STORE_QUANT, 32, float>{});
}

template <class F>
static void dispatch(const aq_sim *s, F &&f) {
  const aq_sim_desc &d = s->d;
  if (d.storage == AQ_STORE_QUANT) {
    const int b = nw32_bucket(d.nwords);
    if (d.dim == 3) dispatch_quant<3, AQ_FLUID>(b, f);
    else if (d.material == AQ_FLUID) dispatch_quant<2, AQ_FLUID>(b, f);
    else dispatch_quant<2, AQ_ELASTIC>(b, f);
  } else if (d.storage == AQ_STORE_F32) {
    if (d.dim == 3) f(Tag<3, AQ_FLUID, AQ_STORE_F32, 4, float>{});
    else if (d.material == AQ_FLUID) f(Tag<2, AQ_FLUID, AQ_STORE_F32, 4, float>{});
    else f(Tag<2, AQ_ELASTIC, AQ_STORE_F32, 4, float>{});
  } else {
    if (d.dim == 3) f(Tag<3, AQ_FLUID, AQ_STORE_F64, 4, double>{});
    else if (d.material == AQ_FLUID) f(Tag<2, AQ_FLUID, AQ_STORE_F64, 4, double>{});
    else f(Tag<2, AQ_ELASTIC, AQ_STORE_F64, 4, double>{});
  }
}

#define TAG_TYPES(tag)                      \
  using TG = decltype(tag);                 \
  constexpr int DIM = TG::DIM;              \
  constexpr int MAT = TG::MAT;              \
  constexpr int ST = TG::ST;                \
  constexpr int NW32 = TG::NW32;            \
  using T = typename TG::T;

template <class T> static const ParamsT<T> &params_of(const aq_sim *s);
template <> const ParamsT<float> &params_of<float>(const aq_sim *s) { return s->pf; }
template <> const ParamsT<double> &params_of<double>(const aq_sim *s) { return s->pd; }

static unsigned pblocks(int64_t n) { return (unsigned)((n + 255) / 256); }
static unsigned gblocks(const aq_sim *s) { return (unsigned)((s->nblocks + 3) / 4); }

static int check_n(const aq_sim *s, const void *state, int64_t n) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  if (n < 0 || n > s->d.capacity) return fail(AQ_ERR_CAPACITY, "particle count exceeds capacity");
  if (n && !state) return fail(AQ_ERR_INVALID, "null state");
  return AQ_OK;
}

static int clear_grid(aq_sim *s, int g, hipStream_t st) {
  if (!s->dirty[g]) return AQ_OK;
  if (s->d.storage == AQ_STORE_F64)
    hipLaunchKernelGGL(k_grid_clear<double>, dim3(gblocks(s)), dim3(256), 0, st, s->nblocks,
                       (double4 *)s->grid[g], s->flags[g], s->act[g]);
  else
    hipLaunchKernelGGL(k_grid_clear<float>, dim3(gblocks(s)), dim3(256), 0, st, s->nblocks,
                       (float4 *)s->grid[g], s->flags[g], s->act[g]);
  LAUNCH_CHECK();
  s->dirty[g] = false;
  return AQ_OK;
}

static int launch_p2g(aq_sim *s, const void *state, int64_t n, int g, hipStream_t st) {
  StateK sk = make_state(s, state);
  if (n == 0) return AQ_OK;
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_p2g<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay,
                       params_of<T>(s), n, (vec4_t<T> *)s->grid[g], s->flags[g]);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

static int launch_grid_update(aq_sim *s, int g, int other, hipStream_t st) {
  void *og = other >= 0 ? s->grid[other] : nullptr;
  uint32_t *of = other >= 0 ? s->flags[other] : nullptr;
  uint32_t *oa = other >= 0 ? s->act[other] : nullptr;
  if (s->d.storage == AQ_STORE_F64) {
    if (s->d.dim == 3)
      hipLaunchKernelGGL((k_grid_update<3, double>), dim3(gblocks(s)), dim3(256), 0, st, s->pd, s->nblocks,
                         (double4 *)s->grid[g], s->flags[g], s->act[g], (double4 *)og, of, oa);
    else
      hipLaunchKernelGGL((k_grid_update<2, double>), dim3(gblocks(s)), dim3(256), 0, st, s->pd, s->nblocks,
                         (double4 *)s->grid[g], s->flags[g], s->act[g], (double4 *)og, of, oa);
  } else {
    if (s->d.dim == 3)
      hipLaunchKernelGGL((k_grid_update<3, float>), dim3(gblocks(s)), dim3(256), 0, st, s->pf, s->nblocks,
                         (float4 *)s->grid[g], s->flags[g], s->act[g], (float4 *)og, of, oa);
    else
      hipLaunchKernelGGL((k_grid_update<2, float>), dim3(gblocks(s)), dim3(256), 0, st, s->pf, s->nblocks,
                         (float4 *)s->grid[g], s->flags[g], s->act[g], (float4 *)og, of, oa);
  }
  LAUNCH_CHECK();
  if (other >= 0) s->dirty[other] = false;
  return AQ_OK;
}

static int launch_g2p(aq_sim *s, void *state, int64_t n, int g, uint32_t step, hipStream_t st) {
  StateK sk = make_state(s, state);
  if (n == 0) return AQ_OK;
  KeysK keys = make_keys(s->d.seed, step);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_g2p<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay, keys,
                       params_of<T>(s), n, (const vec4_t<T> *)s->grid[g]);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_p2g(aq_sim *s, const void *state, int64_t n, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = clear_grid(s, 0, st))) return rc;
  if ((rc = launch_p2g(s, state, n, 0, st))) return rc;
  s->dirty[0] = true;
  s->cur = 0;
  return AQ_OK;
}

extern "C" int aq_grid_op(aq_sim *s, void *stream) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  return launch_grid_update(s, s->cur, -1, (hipStream_t)stream);
}

extern "C" int aq_g2p(aq_sim *s, void *state, int64_t n, uint32_t step, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  return launch_g2p(s, state, n, s->cur, step, (hipStream_t)stream);
}

extern "C" int aq_substep(aq_sim *s, void *state, int64_t n, uint32_t step, void *stream) {
  int rc;
  if ((rc = aq_p2g(s, state, n, stream))) return rc;
  if ((rc = aq_grid_op(s, stream))) return rc;
  return aq_g2p(s, state, n, step, stream);
}

extern "C" int aq_sort(aq_sim *s, void *state, int64_t n, int32_t *ids, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  StateK sk = make_state(s, state);
  hipLaunchKernelGGL(k_fill_u32, dim3((s->nblocks + 1 + 255) / 256), dim3(256), 0, st, s->scount, 0u,
                     (int64_t)s->nblocks + 1);
  LAUNCH_CHECK();
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_sort_keys<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay,
                       params_of<T>(s), n, s->skey, s->srank, s->scount);
  });
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, s->scount, s->soff, s->nblocks);
  LAUNCH_CHECK();
  const int64_t cap = s->d.capacity;
  size_t esz;
  int planes;
  if (s->d.storage == AQ_STORE_QUANT) {
    planes = s->d.nwords;
    esz = 8;
    hipLaunchKernelGGL(k_sort_scatter<uint64_t>, dim3(pblocks(n)), dim3(256), 0, st, (const uint64_t *)state,
                       (uint64_t *)s->tmp_state, planes, cap, n, s->skey, s->srank, s->soff, ids, s->tmp_ids);
  } else if (s->d.storage == AQ_STORE_F32) {
    planes = s->d.ncomp;
    esz = 4;
    hipLaunchKernelGGL(k_sort_scatter<uint32_t>, dim3(pblocks(n)), dim3(256), 0, st, (const uint32_t *)state,
                       (uint32_t *)s->tmp_state, planes, cap, n, s->skey, s->srank, s->soff, ids, s->tmp_ids);
  } else {
    planes = s->d.ncomp;
    esz = 8;
    hipLaunchKernelGGL(k_sort_scatter<uint64_t>, dim3(pblocks(n)), dim3(256), 0, st, (const uint64_t *)state,
                       (uint64_t *)s->tmp_state, planes, cap, n, s->skey, s->srank, s->soff, ids, s->tmp_ids);
  }
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpy2DAsync(state, (size_t)cap * esz, s->tmp_state, (size_t)cap * esz, (size_t)n * esz,
                           (size_t)planes, hipMemcpyDeviceToDevice, st));
  if (ids) HIP_TRY(hipMemcpyAsync(ids, s->tmp_ids, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  return AQ_OK;
}

// Event pairs around launches, summed per category when the call ends.
struct Timer {
  enum { FUSED = 0, GRID = 1, SORT = 2, ENDS = 3, NCAT = 4 };
  hipStream_t st;
  std::vector<hipEvent_t> a, b;
  std::vector<int> cat;
  int launches[NCAT] = {0, 0, 0, 0};
  void begin(int c) {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, st);
    a.push_back(e0);
    b.push_back(e1);
    cat.push_back(c);
    launches[c]++;
  }
  void end() { hipEventRecord(b.back(), st); }
  void collect(float *ms) {
    hipStreamSynchronize(st);
    for (int c = 0; c < NCAT; ++c) ms[c] = 0.0f;
    for (size_t i = 0; i < a.size(); ++i) {
      float t = 0.0f;
      hipEventElapsedTime(&t, a[i], b[i]);
      ms[cat[i]] += t;
      hipEventDestroy(a[i]);
      hipEventDestroy(b[i]);
    }
  }
};

#define TIMED(TM, C, STMT)         \
  do {                             \
    if (TM) (TM)->begin(C);        \
    STMT;                          \
    if (TM) (TM)->end();           \
  } while (0)

static int advance_impl(aq_sim *s, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                        int32_t *ids, hipStream_t st, Timer *tm) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (nsteps < 0) return fail(AQ_ERR_INVALID, "negative step count");
  if (nsteps == 0 || n == 0) return AQ_OK;
  int X = 0, Y = 1;
  if ((rc = clear_grid(s, X, st))) return rc;
  if ((rc = clear_grid(s, Y, st))) return rc;
  if (sort_every > 0) {
    TIMED(tm, Timer::SORT, rc = aq_sort(s, state, n, ids, st));
    if (rc) return rc;
  }
  TIMED(tm, Timer::ENDS, rc = launch_p2g(s, state, n, X, st));
  if (rc) return rc;
  s->dirty[X] = true;
  TIMED(tm, Timer::GRID, rc = launch_grid_update(s, X, -1, st));
  if (rc) return rc;
  for (int t = 0; t + 1 < nsteps; ++t) {
    if (sort_every > 0 && t > 0 && t % sort_every == 0) {
      TIMED(tm, Timer::SORT, rc = aq_sort(s, state, n, ids, st));
      if (rc) return rc;
    }
    StateK sk = make_state(s, state);
    KeysK keys = make_keys(s->d.seed, step0 + (uint32_t)t);
    if (tm) tm->begin(Timer::FUSED);
    dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
      hipLaunchKernelGGL((k_g2p2g<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay, keys,
                         params_of<T>(s), n, (const vec4_t<T> *)s->grid[X], (vec4_t<T> *)s->grid[Y], s->flags[Y]);
    });
    if (tm) tm->end();
    LAUNCH_CHECK();
    s->dirty[Y] = true;
    TIMED(tm, Timer::GRID, rc = launch_grid_update(s, Y, X, st));
    if (rc) return rc;
    std::swap(X, Y);
  }
  TIMED(tm, Timer::ENDS, rc = launch_g2p(s, state, n, X, step0 + (uint32_t)(nsteps - 1), st));
  if (rc) return rc;
  s->cur = X;
  return AQ_OK;
}

extern "C" int aq_advance(aq_sim *s, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                          int32_t *ids, void *stream) {
  return advance_impl(s, state, n, nsteps, step0, sort_every, ids, (hipStream_t)stream, nullptr);
}

__global__ void k_count_active(const uint32_t *__restrict__ act, int32_t nblocks, unsigned long long *out) {
  const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  unsigned long long c = (b < nblocks && act[b]) ? 1ull : 0ull;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

extern "C" int aq_advance_timed(aq_sim *s, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                                int32_t *ids, float *ms, int32_t *launches, int64_t *active_nodes, void *stream) {
  if (!ms || !launches || !active_nodes) return fail(AQ_ERR_INVALID, "null output");
  hipStream_t st = (hipStream_t)stream;
  Timer tm;
  tm.st = st;
  int rc = advance_impl(s, state, n, nsteps, step0, sort_every, ids, st, &tm);
  tm.collect(ms);
  if (rc) return rc;
  for (int c = 0; c < Timer::NCAT; ++c) launches[c] = tm.launches[c];
  // active grid nodes of the last grid (blocks flagged by its grid update)
  unsigned long long *cnt = (unsigned long long *)s->scount;  // sort counters are free here
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)cnt, 0u, (int64_t)2);
  hipLaunchKernelGGL(k_count_active, dim3((s->nblocks + 255) / 256), dim3(256), 0, st, s->act[s->cur], s->nblocks,
                     cnt);
  LAUNCH_CHECK();
  unsigned long long h = 0;
  HIP_TRY(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *active_nodes = (int64_t)h * 64;
  return AQ_OK;
}

extern "C" int aq_update_limits(aq_sim *s, const void *state, int64_t n, float *mxmn, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (!mxmn) return fail(AQ_ERR_INVALID, "null output");
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  StateK sk = make_state(s, state);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_limits<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay, n, mxmn);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_energy(aq_sim *s, const void *state, int64_t n, double *out, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (!out) return fail(AQ_ERR_INVALID, "null output");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)out, 0u, (int64_t)2);
  LAUNCH_CHECK();
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_energy<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, s->lay,
                       params_of<T>(s), n, out);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_sim_grid(aq_sim *s, void **grid, int64_t *nodes) {
  if (!s || !grid || !nodes) return fail(AQ_ERR_INVALID, "null argument");
  *grid = s->grid[s->cur];
  *nodes = s->nodes;
  return AQ_OK;
}
