// libaqmpm: HIP kernels and the C ABI declared in include/aq_mpm.h.
// Built for gfx950 only (see autoquantizer_amd/build.py).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "aq_sim.hpp"
#include "aq_gather.hpp"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

extern "C" const char *aq_last_error(void) { return g_err.c_str(); }
// the error of every translation unit of the library
__attribute__((visibility("hidden"))) int aq_fail(int code, const char *msg) {
  g_err = msg;
  return code;
}
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
      if (!(c.scale > 0.0) || !std::isfinite(c.scale) || !std::isfinite(c.offset))
        return fail(AQ_ERR_INVALID, "fixed-point scale must be positive and finite");
      k.enc_s = (float)(1.0 / c.scale);
      k.dec_s = (float)c.scale;
      k.offs = (float)c.offset;
      // the decode relies on a positive normal f32 scale (dec_fixed)
      if (!(k.dec_s >= 1.17549435e-38f) || !std::isfinite(k.enc_s) || !std::isfinite(k.dec_s))
        return fail(AQ_ERR_INVALID, "fixed-point scale outside the f32 normal range");
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

// ---------------------------------------------------------------------------
// codec kernels
// ---------------------------------------------------------------------------
// Per-thread LDS rows for encoding (odd stride: distinct banks per lane).
template <int NW32> struct Rows {
  static constexpr int RS = NW32 + 1;
};

template <int NW32>
__global__ void __launch_bounds__(256) k_pack(LayoutK lay, KeysK keys, const float *__restrict__ vals,
                                              const float *__restrict__ noise, uint64_t *__restrict__ words,
                                              int64_t n) {
  __shared__ uint32_t rows[256 * Rows<NW32>::RS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t *row = rows + threadIdx.x * Rows<NW32>::RS;
#pragma unroll
  for (int i = 0; i < NW32; ++i) row[i] = 0u;
  uint32_t gmax = 0u;
  for (int c = 0; c < lay.nc; ++c) {
    const CompK &k = lay.c[c];
    const float v = vals[(int64_t)c * n + p];
    if (k.kind == K_FLOAT && k.shared) {
      if (k.gfirst) {
        gmax = 0u;
        for (int j = c; j < lay.nc; ++j) {
          const uint32_t e = f32_exp(vals[(int64_t)j * n + p]);
          gmax = e > gmax ? e : gmax;
          if (lay.c[j].glast) break;
        }
        row_put(row, k.eoff, k.exp_bits, shared_exp_field(gmax, k));
      }
      row_put(row, k.off, k.nbits, enc_shared(v, gmax, k));
    } else {
      float u = 0.0f;
      if (k.kind == K_FIXED && k.dither) u = noise ? noise[(int64_t)c * n + p] : dither_u((uint32_t)p, keys.k[c >> 1], c);
      encode_member(v, k, u, row, false);
    }
  }
  for (int i = 0; i < lay.nw; ++i)
    words[(int64_t)i * n + p] = (uint64_t)row[2 * i] | ((uint64_t)row[2 * i + 1] << 32);
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
// selected component for record idx[t].  Every member of a shared-exponent
// group with a selected member is re-encoded with the group's new exponent,
// the unselected ones from their decoded stored values.
template <int NW32>
__global__ void __launch_bounds__(256) k_store_members(LayoutK lay, KeysK keys, uint32_t mask,
                                                       uint64_t *__restrict__ words, int64_t n,
                                                       const int64_t *__restrict__ idx,
                                                       const float *__restrict__ vals, int64_t count) {
  __shared__ uint32_t rows[256 * Rows<NW32>::RS];
  __shared__ float vvs[256 * (AQ_MAX_COMPS + 1)];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t p = idx[t];
  if (p < 0 || p >= n) return;
  uint32_t *row = rows + threadIdx.x * Rows<NW32>::RS;
  float *vv = vvs + threadIdx.x * (AQ_MAX_COMPS + 1);
  uint32_t r[NW32];
#pragma unroll
  for (int i = 0; i < NW32; ++i) r[i] = 0u;
  for (int i = 0; i < lay.nw; ++i) {
    const uint64_t x = words[(int64_t)i * n + p];
    r[2 * i] = (uint32_t)x;
    r[2 * i + 1] = (uint32_t)(x >> 32);
  }
#pragma unroll
  for (int i = 0; i < NW32; ++i) row[i] = r[i];
  // member values: new where selected, stored otherwise
  int j = 0;
  for (int c = 0; c < lay.nc; ++c) {
    if ((mask >> c) & 1u) {
      vv[c] = vals[(int64_t)j * count + t];
      ++j;
    } else {
      vv[c] = decode_comp(r, lay.c[c]);
    }
  }
  // members to write: selected ones and whole groups touched by them
  uint32_t touched = mask;
  for (int c = 0; c < lay.nc; ++c) {
    const CompK &k = lay.c[c];
    if (k.kind == K_FLOAT && k.shared && ((mask >> c) & 1u)) {
      int g0 = c, g1 = c;
      while (!lay.c[g0].gfirst) --g0;
      while (!lay.c[g1].glast) ++g1;
      for (int q = g0; q <= g1; ++q) touched |= 1u << q;
    }
  }
  uint32_t gmax = 0u;
  for (int c = 0; c < lay.nc; ++c) {
    if (!((touched >> c) & 1u)) continue;
    const CompK &k = lay.c[c];
    if (k.kind == K_FLOAT && k.shared) {
      if (k.gfirst) {
        gmax = 0u;
        for (int q = c; q < lay.nc; ++q) {
          const uint32_t e = f32_exp(vv[q]);
          gmax = e > gmax ? e : gmax;
          if (lay.c[q].glast) break;
        }
        row_set(row, k.eoff, k.exp_bits, shared_exp_field(gmax, k));
      }
      row_set(row, k.off, k.nbits, enc_shared(vv[c], gmax, k));
    } else {
      const float u = (k.kind == K_FIXED && k.dither) ? dither_u((uint32_t)p, keys.k[c >> 1], c) : 0.0f;
      encode_member(vv[c], k, u, row, true);
    }
  }
  for (int i = 0; i < lay.nw; ++i)
    words[(int64_t)i * n + p] = (uint64_t)row[2 * i] | ((uint64_t)row[2 * i + 1] << 32);
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
template <int DIM, class T, bool OP>
__global__ void __launch_bounds__(256) k_gather(ParamsT<T> P, ColliderK K, int32_t nblocks,
                                                const int32_t *__restrict__ item_off,
                                                const uint32_t *__restrict__ near,
                                                const vec4_t<T> *__restrict__ tiles, vec4_t<T> *__restrict__ grid_f,
                                                uint32_t *__restrict__ fflags, vec4_t<T> *__restrict__ grid_v,
                                                uint32_t *__restrict__ act) {
  const int32_t b = __builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6)));
  if (b >= nblocks) return;
  gather_block<DIM, T, OP>(P, K, b, item_off, near, tiles, grid_f, fflags, grid_v, act);
}

// Gather over the block list of the last sort (tiled P2G): gcnt[0] blocks
// of glist (near a particle or active at the sort, and those first reached
// since), then the gcnt[fslot] blocks of flist (flagged by untiled scatters
// since the last gather) that glist does not hold.  A flagged block first
// reached since the sort is appended to glist (at gcnt[3], the append
// cursor), so that later gathers visit it and zero its grid once nothing
// reaches it.  The commit (gcnt[0] = gcnt[3], gcnt[fslot] = 0) waits for
// the next kernel: the next tiled kernel's first thread does it (the
// kernels append into the other counter meanwhile), else k_glist_commit.
// (Committing in the last workgroup to finish costs 2048 serialised
// atomics on one address, +35 us; a kernel of its own after every gather
// 3-4.5 us.)
template <int DIM, class T, bool OP>
__global__ void __launch_bounds__(256) k_gather_list(ParamsT<T> P, ColliderK K, const int32_t *__restrict__ item_off,
                                                     const uint32_t *__restrict__ near,
                                                     const vec4_t<T> *__restrict__ tiles,
                                                     vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags,
                                                     vec4_t<T> *__restrict__ grid_v, uint32_t *__restrict__ act,
                                                     int32_t *__restrict__ glist, const int32_t *__restrict__ flist,
                                                     uint32_t *__restrict__ gcnt, uint32_t *__restrict__ listed,
                                                     int32_t fslot, const uint32_t *__restrict__ bm_read,
                                                     uint32_t *__restrict__ bm_clear) {
  const uint32_t gc = gcnt[0], fc = gcnt[fslot];
  const uint32_t nw = gridDim.x * 4u;
  const int lane = threadIdx.x & 63;
  const uint32_t cap = (uint32_t)(DIM == 3 ? P.nb * P.nb * P.nb : P.nb * P.nb);  // glist's capacity
  uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6)));
  // the wave's first list entry, loaded with the counters (used only when
  // it is below gcnt[0]: slots above may be taking appends meanwhile, so
  // these loads and the appends' stores are relaxed atomics)
  int32_t pre = i < cap ? __hip_atomic_load(glist + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  for (; i < gc + fc; i += nw) {
    int32_t b;
    if (i < gc) {
      b = pre;
    } else {
      b = flist[i - gc];
      if (listed[b]) continue;  // glist holds it
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        listed[b] = 1u;
        __hip_atomic_store(glist + atomicAdd(&gcnt[3], 1u), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    b = __builtin_amdgcn_readfirstlane(b);
    pre = i + nw < cap ? __hip_atomic_load(glist + i + nw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    gather_block<DIM, T, OP, true>(P, K, b, item_off, near, tiles, grid_f, fflags, grid_v, act, bm_read, bm_clear);
  }
}

// The commit of a list gather when no tiled kernel carries it: the blocks
// it appended join glist; its flist counter (slot) is cleared.
__global__ void k_glist_commit(uint32_t *__restrict__ gcnt, int32_t slot) {
  if (threadIdx.x == 0) {
    gcnt[0] = gcnt[3];
    gcnt[slot] = 0u;
  }
}

// The gather list of a sort: every block with particles within one block
// (near) or a grid left from before (act), once (listed[b] marks them);
// slots from gcnt[0], with gcnt[3] counting alike (both zero before).
__global__ void __launch_bounds__(256) k_glist_build(const uint32_t *__restrict__ near,
                                                     const uint32_t *__restrict__ act, uint32_t *__restrict__ listed,
                                                     int32_t *__restrict__ glist, uint32_t *__restrict__ gcnt,
                                                     int32_t m) {
  const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool in = b < m && (near[b] != 0u || act[b] != 0u);
  if (b < m) listed[b] = in ? 1u : 0u;
  const unsigned long long mask = __ballot(in);
  if (mask == 0ull) return;  // uniform over the wavefront
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0u;
  if (lane == leader) {
    base = atomicAdd(&gcnt[0], (uint32_t)__popcll(mask));
    atomicAdd(&gcnt[3], (uint32_t)__popcll(mask));
  }
  base = __shfl(base, leader);
  if (in) glist[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = b;
}

// Grid update in place on the active blocks of grid_v (after a gather without OP).
template <int DIM, class T>
__global__ void __launch_bounds__(256) k_grid_update(ParamsT<T> P, ColliderK K, int32_t nblocks,
                                                     vec4_t<T> *__restrict__ grid_v,
                                                     const uint32_t *__restrict__ act) {
  const int32_t b = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (b >= nblocks || !act[b]) return;
  int bc[3], loc[3], I[3];
  block_coords<DIM>(b, P.nb, bc);
  lane_coords<DIM>(lane, loc);
#pragma unroll
  for (int a = 0; a < 3; ++a) I[a] = bc[a] * BlockShape<DIM>::E + loc[a];
  const int64_t node = (int64_t)b * 64 + lane;
  grid_v[node] = grid_op_node<DIM, T>(P, K, I, grid_v[node]);
}

// Work items from the block counts and offsets of a sort (item_off from the scan).
__global__ void __launch_bounds__(256) k_items_fill(const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ offsets,
                                                    const int32_t *__restrict__ item_off, int4 *__restrict__ items,
                                                    int32_t m) {
  const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= m) return;
  const uint32_t c = counts[b];
  const int32_t base = item_off[b];
  for (uint32_t j = 0; j * (uint32_t)kItem < c; ++j) {
    const uint32_t left = c - j * (uint32_t)kItem;
    items[base + (int32_t)j] =
        make_int4(b, (int)(offsets[b] + j * (uint32_t)kItem), (int)(left < (uint32_t)kItem ? left : (uint32_t)kItem), 0);
  }
}

// near[b]: some block within one block of b (per axis) holds particles, so
// the gather of b may find tiles (the neighbours whose tiles can cover b).
// Also clears, for the new order, b's drift-margin masks of both parities
// (bm[b], bm[m + b]) and the gather list's 16 counters (gcnt).
template <int DIM>
__global__ void __launch_bounds__(256) k_near(const uint32_t *__restrict__ counts, int32_t nb,
                                              uint32_t *__restrict__ near, uint32_t *__restrict__ bm,
                                              uint32_t *__restrict__ gcnt, int32_t m) {
  const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < 16) gcnt[b] = 0u;
  if (b >= m) return;
  bm[b] = 0u;
  bm[m + b] = 0u;
  int bc[3];
  block_coords<DIM>(b, nb, bc);
  uint32_t any = 0u;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz) {
        const int x = bc[0] + dx, y = bc[1] + dy, z = bc[2] + dz;
        if (x < 0 || x >= nb || y < 0 || y >= nb || (DIM == 3 && (z < 0 || z >= nb))) continue;
        const int32_t c = DIM == 3 ? (x * nb + y) * nb + z : x * nb + y;
        any |= counts[c];
      }
  near[b] = any != 0u ? 1u : 0u;
}

__global__ void k_fill_u32(uint32_t *a, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---- sort by grid block of the stencil base --------------------------------
// Keys (grid block of the stencil base), block counts and cell ranks: the
// 64 cells of a block are its stencil bases, and rank = (position among the
// particles of the same cell) << 6 | cell (k_sort_rank turns it into the
// position in the block).  Lanes of a wavefront with equal keys add to the
// block count with one atomic (particles are mostly sorted already, so a
// wavefront holds few distinct keys); after four distinct keys the rest add
// one each.
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_sort_keys(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                                   uint32_t *__restrict__ key, uint32_t *__restrict__ rank,
                                                   uint32_t *__restrict__ counts, uint32_t *__restrict__ ccounts) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < n;
  uint32_t k = 0u, cell = 0u;
  if (valid) {
    T x[DIM];
    if constexpr (ST == AQ_STORE_QUANT) {
      uint32_t *row = rows + threadIdx.x * RowsOf<ST, NW32>::RS;
      const uint64_t *w = (const uint64_t *)s.ptr;
#pragma unroll
      for (int i = 0; i < NW32 / 2; ++i) {
        if (i < s.nw && ((s.g2p_mask >> i) & 1u)) {
          const uint64_t z = w[(int64_t)i * s.cap + p];
          row[2 * i] = (uint32_t)z;
          row[2 * i + 1] = (uint32_t)(z >> 32);
        }
      }
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        if constexpr (IsF64<T>::value)
          x[a] = decode_row_f64(row, comp_at(cod, a), cod.cd + 4 * a);  // f64 compute
        else
          x[a] = (T)decode_row(row, comp_at(cod, a));
      }
    } else {
      const T *vals = (const T *)s.ptr;
#pragma unroll
      for (int a = 0; a < DIM; ++a) x[a] = vals[(int64_t)a * s.cap + p];
    }
    int base[3] = {0, 0, 0};
    T fx, w[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) base[a] = stencil_1d<T>(x[a], P, fx, w);
    k = (uint32_t)block_of_base<DIM>(base, P.nb);
    constexpr int E = BlockShape<DIM>::E, LG = BlockShape<DIM>::LOG;
    cell = DIM == 3 ? (uint32_t)(((base[0] & (E - 1)) << (2 * LG)) | ((base[1] & (E - 1)) << LG) | (base[2] & (E - 1)))
                    : (uint32_t)(((base[0] & (E - 1)) << LG) | (base[1] & (E - 1)));
    key[p] = k;
  }
  uint64_t pending = __ballot(valid);
  for (int round = 0; round < 4 && pending; ++round) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t k0 = __shfl(k, leader);
    const uint64_t same = __ballot(k == k0) & pending;
    if (lane == leader) atomicAdd(counts + k0, (uint32_t)__popcll(same));
    pending &= ~same;
  }
  if ((pending >> lane) & 1ull) atomicAdd(counts + k, 1u);
  if (valid) rank[p] = (atomicAdd(ccounts + (int64_t)k * 64 + cell, 1u) << 6) | cell;
}

// Position of each particle in its block, layer by layer: layer r holds the
// r-th particle of every cell that has more than r, in cell order, so that
// the 64 lanes of a wavefront of the tiled kernels mostly hold particles of
// distinct cells (their P2G atomics and G2P reads then hit distinct LDS
// slots; in particle order within a cell, ~8 lanes would share each slot).
// pos = sum_c min(count_c, r) + #{c < cell : count_c > r}.
__global__ void __launch_bounds__(256) k_sort_rank(const uint32_t *__restrict__ key, uint32_t *__restrict__ rank,
                                                   const uint32_t *__restrict__ ccounts, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t rc = rank[p], r = rc >> 6, cell = rc & 63u;
  const uint4 *cc = (const uint4 *)(ccounts + (int64_t)key[p] * 64);
  uint32_t pos = 0u;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint4 c4 = cc[i];
    const uint32_t c[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) pos += (c[j] < r ? c[j] : r) + ((uint32_t)(4 * i + j) < cell && c[j] > r ? 1u : 0u);
  }
  rank[p] = pos;
}

// Exclusive scans of counts[0..m) into off[0..m] and of the work items per
// block, ceil(counts / kItem), into ioff[0..m] (totals at [m]): per-chunk scans
// of 4096 (k_scan_local), a scan of the chunk totals (k_scan_top) and the
// chunk prefix added back (k_scan_add).
// Work items of a block with c particles (k_items_fill writes exactly these).
__device__ __forceinline__ uint32_t items_of(uint32_t c) { return (c + (uint32_t)kItem - 1u) / (uint32_t)kItem; }

__global__ void __launch_bounds__(256) k_scan_local(const uint32_t *__restrict__ counts, uint32_t *__restrict__ off,
                                                    int32_t *__restrict__ ioff, uint32_t *__restrict__ part,
                                                    int32_t m) {
  __shared__ uint32_t sa[256], sb[256];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 4096 + t * 16;
  uint32_t c[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) c[i] = base + i < m ? counts[base + i] : 0u;
  uint32_t a = 0u, b = 0u;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    a += c[i];
    b += items_of(c[i]);
  }
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t x = t >= o ? sa[t - o] : 0u, y = t >= o ? sb[t - o] : 0u;
    __syncthreads();
    sa[t] += x;
    sb[t] += y;
    __syncthreads();
  }
  uint32_t ra = sa[t] - a, rb = sb[t] - b;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (base + i < m) {
      off[base + i] = ra;
      ioff[base + i] = (int32_t)rb;
    }
    ra += c[i];
    rb += items_of(c[i]);
  }
  if (t == 255) {
    part[2 * blockIdx.x] = sa[255];
    part[2 * blockIdx.x + 1] = sb[255];
  }
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t *__restrict__ part, int32_t g, uint32_t *__restrict__ off,
                                                   int32_t *__restrict__ ioff, int32_t m) {
  __shared__ uint32_t sa[1024], sb[1024];
  const int t = threadIdx.x;
  const int per = (g + 1023) / 1024;
  const int lo = t * per, hi = lo + per < g ? lo + per : g;
  uint32_t a = 0u, b = 0u;
  for (int i = lo; i < hi; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t x = t >= o ? sa[t - o] : 0u, y = t >= o ? sb[t - o] : 0u;
    __syncthreads();
    sa[t] += x;
    sb[t] += y;
    __syncthreads();
  }
  uint32_t ra = sa[t] - a, rb = sb[t] - b;
  for (int i = lo; i < hi; ++i) {
    const uint32_t pa = part[2 * i], pb = part[2 * i + 1];
    part[2 * i] = ra;
    part[2 * i + 1] = rb;
    ra += pa;
    rb += pb;
  }
  if (t == 1023) {
    off[m] = sa[1023];
    ioff[m] = (int32_t)sb[1023];
  }
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t *__restrict__ off, int32_t *__restrict__ ioff,
                                                  const uint32_t *__restrict__ part, int32_t m) {
  const uint32_t pa = part[2 * blockIdx.x], pb = part[2 * blockIdx.x + 1];
  const int64_t base = (int64_t)blockIdx.x * 4096;
  for (int i = threadIdx.x; i < 4096; i += 256) {
    if (base + i < m) {
      off[base + i] += pa;
      ioff[base + i] += (int32_t)pb;
    }
  }
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
// Per-component max / min of the state: grid-stride blocks write their
// partial max / min to part[block][2 NC] (no atomics: many blocks updating
// the same few addresses serialise); k_limits_final folds them into mxmn.
constexpr int kLimitBlocks = 512;

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_limits(StateK s, CodecK cod, int64_t n, float *__restrict__ part) {
  constexpr int NC = Comps<DIM, MAT>::NC;
  __shared__ float smx[NC][4], smn[NC][4];
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  float mx[NC], mn[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    mx[c] = -INFINITY;
    mn[c] = INFINITY;
  }
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    T v[NC];
    load_state<DIM, ST, NW32, NC, T>(s, cod, p, 0xFFFFFFFFu, rows + threadIdx.x * RowsOf<ST, NW32>::RS, v);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      mx[c] = fmaxf(mx[c], (float)v[c]);
      mn[c] = fminf(mn[c], (float)v[c]);
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
    const float a = fmaxf(fmaxf(smx[c][0], smx[c][1]), fmaxf(smx[c][2], smx[c][3]));
    const float b = fminf(fminf(smn[c][0], smn[c][1]), fminf(smn[c][2], smn[c][3]));
    part[(int64_t)blockIdx.x * 2 * NC + c] = a;
    part[(int64_t)blockIdx.x * 2 * NC + NC + c] = b;
  }
}

// mxmn[c] = max(mxmn[c], partial maxima), mxmn[NC + c] = min(...): one
// workgroup, wavefront w folding components w, w + 4, ...
__global__ void __launch_bounds__(256) k_limits_final(const float *__restrict__ part, int32_t nparts, int32_t nc,
                                                      float *__restrict__ mxmn) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = wv; c < nc; c += 4) {
    float a = -INFINITY, b = INFINITY;
    for (int i = lane; i < nparts; i += 64) {
      a = fmaxf(a, part[(int64_t)i * 2 * nc + c]);
      b = fminf(b, part[(int64_t)i * 2 * nc + nc + c]);
    }
    for (int o = 32; o > 0; o >>= 1) {
      a = fmaxf(a, __shfl_xor(a, o));
      b = fminf(b, __shfl_xor(b, o));
    }
    if (lane == 0) {
      mxmn[c] = fmaxf(mxmn[c], a);
      mxmn[nc + c] = fminf(mxmn[nc + c], b);
    }
  }
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_energy(StateK s, CodecK cod, ParamsT<T> P, int64_t n, double *__restrict__ out) {
  constexpr int NC = Comps<DIM, MAT>::NC;
  using CI = Comps<DIM, MAT>;
  __shared__ double part[8];
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (p < n) {
    T v[NC];
    load_state<DIM, ST, NW32, NC, T>(s, cod, p, 0xFFFFFFFFu, rows + threadIdx.x * RowsOf<ST, NW32>::RS, v);
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
// Whether the record's fields, in the encoder's order (components in order,
// a shared group's exponent before its first member, an own exponent before
// its digits), lie at non-decreasing bit offsets: then records are written
// as a bit stream from registers (StreamSink).
static bool stream_ordered(const LayoutK &lay) {
  uint32_t pos = 0;
  auto put = [&](uint32_t off, uint32_t nb) {
    if (off < pos) return false;
    pos = off + nb;
    return pos <= (uint32_t)lay.nw * 64u;
  };
  for (int c = 0; c < lay.nc; ++c) {
    const CompK &k = lay.c[c];
    if (k.kind == K_FLOAT && (!k.shared || k.gfirst))
      if (!put(k.eoff, k.exp_bits)) return false;
    if (!put(k.off, k.kind == K_F32 ? 32u : k.nbits)) return false;
  }
  return true;
}


static int check_desc(const aq_sim_desc *d) {
  if (!d) return fail(AQ_ERR_INVALID, "null descriptor");
  if (d->dim != 2 && d->dim != 3) return fail(AQ_ERR_INVALID, "dim must be 2 or 3");
  if (d->material != AQ_FLUID && d->material != AQ_ELASTIC) return fail(AQ_ERR_INVALID, "bad material");

  const int e = d->dim == 3 ? 4 : 8;
  if (d->n_grid < 8 || d->n_grid % e) return fail(AQ_ERR_INVALID, "n_grid must be a multiple of the block edge");
  if (d->dim == 3 && d->n_grid > 1024) return fail(AQ_ERR_INVALID, "n_grid too large");
  if (d->dim == 2 && d->n_grid > 32768) return fail(AQ_ERR_INVALID, "n_grid too large");
  if (d->storage < AQ_STORE_QUANT || d->storage > AQ_STORE_F64) return fail(AQ_ERR_INVALID, "bad storage");
  const int nc = 2 * d->dim + d->dim * d->dim + (d->material == AQ_FLUID ? 1 : d->dim * d->dim);
  if (d->ncomp != nc) return fail(AQ_ERR_INVALID, "ncomp does not match dim/material");
  if (d->capacity < 0 || d->capacity > 0x7FFFFFFFll) return fail(AQ_ERR_INVALID, "capacity out of range");
  if (d->boundary != AQ_SEPARATING && d->boundary != AQ_STICKY) return fail(AQ_ERR_INVALID, "bad boundary");
  if (d->bound < 0 || 2 * d->bound >= d->n_grid) return fail(AQ_ERR_INVALID, "bad wall thickness");
  if (d->axis_bounds != 0 && d->axis_bounds != 1) return fail(AQ_ERR_INVALID, "axis_bounds must be 0 or 1");
  if (d->axis_bounds)
    for (int a = 0; a < d->dim; ++a)
      if (d->lb[a] < 0 || d->lb[a] >= d->ub[a] || d->ub[a] > d->n_grid)
        return fail(AQ_ERR_INVALID, "per-axis walls need 0 <= lb < ub <= n_grid");
  if (d->bound_clip != 0 && d->bound_clip != 1) return fail(AQ_ERR_INVALID, "bound_clip must be 0 or 1");
  if (d->compute_f64 != 0 && d->compute_f64 != 1) return fail(AQ_ERR_INVALID, "compute_f64 must be 0 or 1");
  if (d->compute_f64 && (d->storage != AQ_STORE_QUANT || d->dim != 2))
    return fail(AQ_ERR_INVALID, "compute_f64 applies to quantized storage in 2D");
  return AQ_OK;
}

static int64_t blocks_of(const aq_sim_desc *d) {
  const int64_t n = d->n_grid;
  return (d->dim == 3 ? n * n * n : n * n) / 64;
}

// Work items of a sort: at most ceil(count_b / kItem) per block, so at most
// capacity / kItem plus one per non-empty block.
static int32_t max_items_of(const aq_sim_desc *d) {
  const int64_t nb = blocks_of(d);
  const int64_t ne = d->capacity < nb ? d->capacity : nb;
  return (int32_t)(d->capacity / kItem + ne + 1);
}

static int64_t tile_slots(const aq_sim_desc *d) { return d->dim == 3 ? Tile<3>::STRIDE : Tile<2>::STRIDE; }

enum {
  WS_GRID_V, WS_GRID_F, WS_FLAGS, WS_GLIST, WS_KEY, WS_RANK, WS_COUNT, WS_CCOUNT, WS_OFF, WS_PART, WS_TMP_STATE, WS_TMP_IDS,
  WS_ITEMS, WS_ITEM_OFF, WS_NEAR, WS_TILES, WS_LIMITS, WS_LAYOUT, WS_N
};

static void ws_plan(const aq_sim_desc *d, int64_t *offs /*[WS_N]*/, int64_t *total) {
  const int64_t nblocks = blocks_of(d);
  const int64_t nodes = nblocks * 64;
  const int64_t node_bytes = f64_compute(*d) ? 32 : 16;
  const int64_t cap = d->capacity;
  const int64_t state_bytes = d->storage == AQ_STORE_QUANT ? cap * 8 * d->nwords
                              : cap * d->ncomp * (d->storage == AQ_STORE_F64 ? 8 : 4);
  const int64_t chunks = (nblocks + 1 + 4095) / 4096;
  int64_t sz[WS_N];
  sz[WS_GRID_V] = nodes * node_bytes;
  sz[WS_GRID_F] = nodes * node_bytes;
  sz[WS_FLAGS] = nblocks * 4 * 4;  // fflags, act, two drift-margin flag arrays
  sz[WS_GLIST] = 64 + nblocks * 4 * 3;  // counters; listed flags, block list, flagged-block list
  sz[WS_KEY] = cap * 4;
  sz[WS_RANK] = cap * 4;
  sz[WS_COUNT] = (nblocks + 1) * 4;
  sz[WS_CCOUNT] = nodes * 4;  // particles per cell (64 cells per block)
  sz[WS_OFF] = (nblocks + 1) * 4;
  sz[WS_PART] = chunks * 8;
  sz[WS_TMP_STATE] = state_bytes;
  sz[WS_TMP_IDS] = cap * 4;
  sz[WS_ITEMS] = (int64_t)max_items_of(d) * 16;
  sz[WS_ITEM_OFF] = (nblocks + 1) * 4;
  sz[WS_NEAR] = nblocks * 4;
  sz[WS_TILES] = (int64_t)max_items_of(d) * tile_slots(d) * node_bytes;
  sz[WS_LIMITS] = (int64_t)kLimitBlocks * 2 * AQ_MAX_COMPS * 4;  // partial max / min of k_limits
  sz[WS_LAYOUT] = AQ_MAX_COMPS * (int64_t)(sizeof(CompK) + 4 * sizeof(double));  // CompK table, f64 factors
  int64_t o = 0;
  for (int i = 0; i < WS_N; ++i) {
    offs[i] = o;
    o = align256(o + sz[i]);
  }
  *total = o;
}

extern "C" int64_t aq_sim_workspace_bytes(const aq_sim_desc *d) {
  if (check_desc(d)) return -1;
  int64_t offs[WS_N], total;
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
  P.boundary = d->boundary;
  P.clip = d->bound_clip;
  for (int a = 0; a < 3; ++a) {
    P.lb[a] = d->axis_bounds && a < d->dim ? d->lb[a] : d->bound;
    P.ub[a] = d->axis_bounds && a < d->dim ? d->ub[a] : d->n_grid - d->bound;
    P.xlo[a] = (T)(P.lb[a] * d->dx);  // self.lb[k] * self.dx, a Python double folded into the kernel
    P.xhi[a] = (T)(P.ub[a] * d->dx);
  }
  return P;
}

extern "C" int aq_sim_create(const aq_sim_desc *d, void *ws, int64_t ws_bytes, aq_sim **out) {
  int rc = check_desc(d);
  if (rc) return rc;
  // the tiled P2G packs each node's mass as a non-negative fixed-point sum
  if (!(d->p_mass > 0.0) || !std::isfinite(d->p_mass))
    return fail(AQ_ERR_INVALID, "particle mass must be positive and finite");
  if (!out) return fail(AQ_ERR_INVALID, "null output");
  int64_t offs[WS_N], total;
  ws_plan(d, offs, &total);
  if (!ws || ws_bytes < total) return fail(AQ_ERR_CAPACITY, "workspace too small");
  aq_sim *s = new aq_sim();
  s->d = *d;
  s->dither_any = false;
  s->stream_order = false;
  if (d->storage == AQ_STORE_QUANT) {
    rc = build_layout(d->comps, d->ncomp, d->nwords, &s->lay);
    if (rc) {
      delete s;
      return rc;
    }
    for (int c = 0; c < d->ncomp; ++c) s->dither_any = s->dither_any || s->lay.c[c].dither;
    s->stream_order = stream_ordered(s->lay);
  } else {
    std::memset(&s->lay, 0, sizeof(s->lay));
  }
  s->pf = make_params<float>(d);
  s->pd = make_params<double>(d);
  s->nblocks = (int32_t)blocks_of(d);
  s->nodes = (int64_t)s->nblocks * 64;
  s->node_bytes = f64_compute(*d) ? 32 : 16;
  char *base = (char *)ws;
  s->grid_v = base + offs[WS_GRID_V];
  s->grid_f = base + offs[WS_GRID_F];
  s->fflags = (uint32_t *)(base + offs[WS_FLAGS]);
  s->act = s->fflags + s->nblocks;
  s->bmargin = s->act + s->nblocks;
  s->gcnt = (uint32_t *)(base + offs[WS_GLIST]);
  s->listed = s->gcnt + 16;
  s->glist = (int32_t *)(s->listed + s->nblocks);
  s->flist = s->glist + s->nblocks;
  s->skey = (uint32_t *)(base + offs[WS_KEY]);
  s->srank = (uint32_t *)(base + offs[WS_RANK]);
  s->scount = (uint32_t *)(base + offs[WS_COUNT]);
  s->ccount = (uint32_t *)(base + offs[WS_CCOUNT]);
  s->soff = (uint32_t *)(base + offs[WS_OFF]);
  s->spart = (uint32_t *)(base + offs[WS_PART]);
  s->tmp_state = base + offs[WS_TMP_STATE];
  s->tmp_ids = (int32_t *)(base + offs[WS_TMP_IDS]);
  s->items = (int4 *)(base + offs[WS_ITEMS]);
  s->item_off = (int32_t *)(base + offs[WS_ITEM_OFF]);
  s->tiles = base + offs[WS_TILES];
  s->near = (uint32_t *)(base + offs[WS_NEAR]);
  s->lim_part = (float *)(base + offs[WS_LIMITS]);
  s->lay_dev = (CompK *)(base + offs[WS_LAYOUT]);
  s->lay_f64 = (double *)(s->lay_dev + AQ_MAX_COMPS);
  double lf[AQ_MAX_COMPS * 4];
  std::memset(lf, 0, sizeof(lf));
  if (d->storage == AQ_STORE_QUANT)
    for (int c = 0; c < d->ncomp; ++c) {
      const double sc = d->comps[c].scale;
      lf[4 * c + 0] = sc != 0.0 ? 1.0 / sc : 0.0;  // float_to_custom_int's 1 / scale, in the compute type
      lf[4 * c + 1] = sc;
      lf[4 * c + 2] = d->comps[c].offset;
    }
  s->max_items = max_items_of(d);
  s->have_items = false;
  s->items_state = nullptr;
  s->items_n = -1;
  s->ws_bytes = ws_bytes;
  // grids and block flags start at zero; the layout table goes to the device
  hipError_t e = hipMemset(ws, 0, (size_t)offs[WS_KEY]);
  if (e == hipSuccess) e = hipMemcpy(s->lay_dev, s->lay.c, sizeof(s->lay.c), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->lay_f64, lf, sizeof(lf), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    delete s;
    return fail(AQ_ERR_HIP, std::string("workspace set-up: ") + hipGetErrorString(e));
  }
  *out = s;
  return AQ_OK;
}

extern "C" int aq_sim_destroy(aq_sim *s) {
  if (s && s->jit_mod) (void)hipModuleUnload(s->jit_mod);
  delete s;
  return AQ_OK;
}

StateK make_state(const aq_sim *s, const void *state, const int32_t *ids) {
  StateK k;
  k.ptr = (void *)state;
  k.cap = s->d.capacity;
  k.ids = ids;
  k.nw = s->d.nwords;
  k.dither_any = s->dither_any ? 1 : 0;
  k.pad = 0;
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

CodecK make_codec(const aq_sim *s, uint32_t step) {
  CodecK c;
  c.c = s->lay_dev;
  c.nc = s->lay.nc;
  c.nw = s->lay.nw;
  c.shared_mask = c.gfirst_mask = c.glast_mask = 0u;
  if (s->d.storage == AQ_STORE_QUANT) {
    for (int i = 0; i < s->lay.nc; ++i) {
      const CompK &k = s->lay.c[i];
      if (k.kind == K_FLOAT && k.shared) {
        c.shared_mask |= 1u << i;
        if (k.gfirst) c.gfirst_mask |= 1u << i;
        if (k.glast) c.glast_mask |= 1u << i;
      }
    }
  }
  c.key_base = key_base(s->d.seed, step);
  c.stream = s->stream_order ? 1u : 0u;
  c.pad = 0u;
  c.wraps = s->wraps;
  c.cd = s->d.storage == AQ_STORE_QUANT && s->d.compute_f64 ? s->lay_f64 : nullptr;
  return c;
}

// ---- run-time specialisation of the fused kernel to one record layout -----
// The reference compiles every kernel for its quantization scheme (Taichi's
// JIT bakes bit offsets and scales into the code).  Here the fused G2P/P2G
// kernel is compiled once more with hipRTC for a simulation's layout: the
// generated source defines the descriptors and masks as constants and
// includes aq_tiled.hpp, so kinds, offsets, widths and scales fold into the
// code (no descriptor loads, no kind dispatch, constant shifts).  The code
// object is cached on disk under a hash of the source, the headers and the
// options.

static const char *const kJitOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
static const char *const kJitHeaders[] = {"aq_tiled.hpp", "aq_mpm_kernels.hpp", "aq_codec.hpp"};
// 3 waves per SIMD: as fast as 4 on the 3D fluid record (measured), without
// the register spills 4 waves force on the specialised code.
constexpr int kJitWaves = 3;

static uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

static bool read_file(const std::string &path, std::vector<char> &out) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  out.clear();
  char buf[65536];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + r);
  std::fclose(f);
  return true;
}

// A host-only view of a descriptor's layout (no device memory): what the
// source generator reads.
static int jit_layout(const aq_sim_desc *d, aq_sim &v) {
  if (!d) return fail(AQ_ERR_INVALID, "null descriptor");
  if (d->storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "specialisation applies to quantized storage");
  int rc = check_desc(d);
  if (rc) return rc;
  v.d = *d;
  if ((rc = build_layout(d->comps, d->ncomp, d->nwords, &v.lay))) return rc;
  v.lay_dev = nullptr;
  v.dither_any = false;
  for (int c = 0; c < d->ncomp; ++c) v.dither_any = v.dither_any || v.lay.c[c].dither;
  v.stream_order = stream_ordered(v.lay);
  return AQ_OK;
}

static int jit_source(const aq_sim &v, std::string &src) {
  const CodecK c = make_codec(&v, 0u);
  const StateK k = make_state(&v, nullptr, nullptr);
  char buf[1024];
  // hipRTC has no C headers: the fixed-width types come from its runtime header
  src = "// generated: k_tiled<G2P+P2G> specialised to one record layout\n";
  for (const char *t : {"int8_t", "uint8_t", "int16_t", "uint16_t", "int32_t", "uint32_t", "int64_t", "uint64_t"})
    src += std::string("typedef __hip_internal::") + t + " " + t + ";\n";
  src += "#define AQ_JIT_LAYOUT 1\n#define AQ_JIT_COMPS {";
  for (int i = 0; i < v.lay.nc; ++i) {
    const CompK &q = v.lay.c[i];
    if (!std::isfinite(q.enc_s) || !std::isfinite(q.dec_s) || !std::isfinite(q.offs))
      return fail(AQ_ERR_INVALID, "non-finite component scale");
    std::snprintf(buf, sizeof buf, "%s{%u, %u, %u, %u, %u, %u, %u, %u, %u, %u, %uu, %a, %a, %a, 0u}", i ? ", " : "",
                  q.kind, q.nbits, q.is_signed, q.dither, q.exp_bits, q.shared, q.gfirst, q.glast, q.off, q.eoff,
                  q.conv, (double)q.enc_s, (double)q.dec_s, (double)q.offs);
    src += buf;
  }
  std::snprintf(buf, sizeof buf,
                "}\n#define AQ_JIT_SHARED %uu\n#define AQ_JIT_GFIRST %uu\n#define AQ_JIT_GLAST %uu\n"
                "#define AQ_JIT_NW %d\n#define AQ_JIT_G2P_MASK %uu\n#include \"aq_tiled.hpp\"\n",
                c.shared_mask, c.gfirst_mask, c.glast_mask, v.lay.nw, k.g2p_mask);
  src += buf;
  // the fused kernel, and the P2G-only and G2P-only kernels that open and
  // close a run of substeps
  static const char *const kinds[3][3] = {
      {"aq_jit_tiled", "true", "true"}, {"aq_jit_p2g", "false", "true"}, {"aq_jit_g2p", "true", "false"}};
  for (const auto &kd : kinds) {
    std::snprintf(buf, sizeof buf,
                  "extern \"C\" __global__ void __launch_bounds__(%d) __attribute__((amdgpu_waves_per_eu(%d)))\n"
                  "%s(aq::StateK s, aq::CodecK cod, aq::ParamsT<float> P, int64_t n, const int4 *items,\n"
                  "   const int32_t *nitems, const float4 *grid_v, float4 *tiles, float4 *grid_f,\n"
                  "   uint32_t *fflags, int32_t *flist, uint32_t *gcnt, uint32_t ctl, uint32_t *bmargin) {\n"
                  "  aq::tiled_body<%d, %d, AQ_STORE_QUANT, %d, float, %s, %s>(s, cod, P, n, items, nitems, grid_v,\n"
                  "      tiles, grid_f, fflags, flist, gcnt, ctl, bmargin);\n}\n",
                  kItem, kJitWaves, kd[0], v.d.dim, v.d.material, nw32_bucket(v.d.nwords), kd[1], kd[2]);
    src += buf;
  }
  return AQ_OK;
}

// Code object of the specialised kernel: from the cache when present, else
// compiled (and cached when cache_dir is given).  path (optional) receives
// the cache file name.
static int jit_code(const aq_sim &v, const char *src_dir, const char *include_dir, const char *cache_dir,
                    std::vector<char> &code, std::string *path) {
  if (!src_dir || !include_dir) return fail(AQ_ERR_INVALID, "null source directory");
  std::string src;
  int rc = jit_source(v, src);
  if (rc) return rc;
  uint64_t h = fnv1a(1469598103934665603ull, src.data(), src.size());
  for (const char *o : kJitOptions) h = fnv1a(h, o, std::strlen(o) + 1);
  std::vector<char> text;
  for (const char *f : kJitHeaders) {
    if (!read_file(std::string(src_dir) + "/" + f, text)) return fail(AQ_ERR_INVALID, "kernel header not found");
    h = fnv1a(h, text.data(), text.size());
  }
  if (!read_file(std::string(include_dir) + "/aq_mpm.h", text)) return fail(AQ_ERR_INVALID, "aq_mpm.h not found");
  h = fnv1a(h, text.data(), text.size());
  // a different compiler gives a different code object
  int rtc_major = 0, rtc_minor = 0;
  if (hiprtcVersion(&rtc_major, &rtc_minor) != HIPRTC_SUCCESS) return fail(AQ_ERR_HIP, "hiprtcVersion failed");
  const int rtc_ver[2] = {rtc_major, rtc_minor};
  h = fnv1a(h, rtc_ver, sizeof rtc_ver);
  const int hip_ver = HIP_VERSION;
  h = fnv1a(h, &hip_ver, sizeof hip_ver);
  std::string file;
  if (cache_dir) {
    // code objects found here are loaded and run: only a directory that no
    // other user can write to is trusted
    struct stat cs;
    if (::stat(cache_dir, &cs) != 0 || !S_ISDIR(cs.st_mode))
      return fail(AQ_ERR_INVALID, "JIT cache directory does not exist");
    if (cs.st_uid != ::geteuid()) return fail(AQ_ERR_INVALID, "JIT cache directory is not owned by the current user");
    if (cs.st_mode & (S_IWGRP | S_IWOTH))
      return fail(AQ_ERR_INVALID, "JIT cache directory is writable by other users");
    char name[64];
    std::snprintf(name, sizeof name, "/aqjit_%016llx.co", (unsigned long long)h);
    file = std::string(cache_dir) + name;
    if (path) *path = file;
    if (read_file(file, code) && !code.empty()) return AQ_OK;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "aq_jit_tiled.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return fail(AQ_ERR_HIP, "hiprtcCreateProgram failed");
  const std::string i1 = std::string("-I") + src_dir, i2 = std::string("-I") + include_dir;
  std::vector<const char *> opts(std::begin(kJitOptions), std::end(kJitOptions));
  opts.push_back(i1.c_str());
  opts.push_back(i2.c_str());
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n) {
      log.resize(n);
      (void)hiprtcGetProgramLog(prog, &log[0]);
    }
    (void)hiprtcDestroyProgram(&prog);
    return fail(AQ_ERR_HIP, ("hipRTC compile of the specialised kernel failed: " + log.substr(0, 1500)).c_str());
  }
  size_t n = 0;
  if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
    (void)hiprtcDestroyProgram(&prog);
    return fail(AQ_ERR_HIP, "hipRTC produced no code");
  }
  code.resize(n);
  const hiprtcResult gr = hiprtcGetCode(prog, code.data());
  (void)hiprtcDestroyProgram(&prog);
  if (gr != HIPRTC_SUCCESS) return fail(AQ_ERR_HIP, "hiprtcGetCode failed");
  if (cache_dir) {  // write then rename: concurrent builders never see a partial file
    const std::string tmp = file + ".tmp" + std::to_string((unsigned long long)(uintptr_t)&code);
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (f) {
      const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
      std::fclose(f);
      if (!ok || std::rename(tmp.c_str(), file.c_str()) != 0) std::remove(tmp.c_str());
    }
  }
  return AQ_OK;
}

extern "C" int aq_jit_build(const aq_sim_desc *desc, const char *src_dir, const char *include_dir,
                            const char *cache_dir, char *path, int64_t path_cap) {
  if (!cache_dir || !path || path_cap < 1) return fail(AQ_ERR_INVALID, "cache directory and path buffer needed");
  aq_sim v;
  int rc = jit_layout(desc, v);
  if (rc) return rc;
  std::vector<char> code;
  std::string file;
  if ((rc = jit_code(v, src_dir, include_dir, cache_dir, code, &file))) return rc;
  if ((int64_t)file.size() + 1 > path_cap) return fail(AQ_ERR_CAPACITY, "path buffer too small");
  std::memcpy(path, file.c_str(), file.size() + 1);
  return AQ_OK;
}

extern "C" int aq_sim_specialize(aq_sim *s, const char *src_dir, const char *include_dir, const char *cache_dir) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  if (s->jit_mod) {
    (void)hipModuleUnload(s->jit_mod);
    s->jit_mod = nullptr;
    s->jit_fn = s->jit_fn_p2g = s->jit_fn_g2p = nullptr;
  }
  if (!src_dir) return AQ_OK;  // back to the generic kernel
  if (s->d.storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "specialisation applies to quantized storage");
  if (s->d.compute_f64) return fail(AQ_ERR_INVALID, "the specialised kernel computes in f32");
  std::vector<char> code;
  int rc = jit_code(*s, src_dir, include_dir, cache_dir, code, nullptr);
  if (rc) return rc;
  hipModule_t mod;
  HIP_TRY(hipModuleLoadData(&mod, code.data()));
  hipFunction_t fn[3];
  const char *const names[3] = {"aq_jit_tiled", "aq_jit_p2g", "aq_jit_g2p"};
  for (int i = 0; i < 3; ++i) {
    const hipError_t e = hipModuleGetFunction(&fn[i], mod, names[i]);
    if (e != hipSuccess) {
      (void)hipModuleUnload(mod);
      return fail(AQ_ERR_HIP, hipGetErrorString(e));
    }
  }
  s->jit_mod = mod;
  s->jit_fn = fn[0];
  s->jit_fn_p2g = fn[1];
  s->jit_fn_g2p = fn[2];
  return AQ_OK;
}

// The collider at substep `step` (aq_collider's schedule).
ColliderK collider_at(const aq_sim *s, uint32_t step) {
  ColliderK K{};
  if (!s->has_collider) return K;
  const aq_collider &c = s->col;
  const int64_t sub = (int64_t)step - (int64_t)c.first_step;  // the collider's substep
  int64_t k = sub - c.lag;
  if (k < 0) k = 0;
  const double theta = c.omega * s->d.dt * (double)k;
  const double height = c.height_rate * s->d.dt * (double)k;
  K.sdf = c.sdf;
  K.normal = c.normal;
  for (int a = 0; a < 3; ++a) {
    K.size[a] = c.size[a];
    K.offset[a] = c.offset[a];
    K.center[a] = c.center[a];
  }
  K.on = 1;
  K.yshift = (int32_t)(height / s->d.dx + 0.5);
  K.from_rot = c.velocity_from_rotated;
  K.cos_t = std::cos(theta);
  K.sin_t = std::sin(theta);
  K.omega = sub >= 1 ? c.omega : 0.0;
  return K;
}

int settle_list(aq_sim *s, hipStream_t st) {
  if (s->commit_pending) {
    hipLaunchKernelGGL(k_glist_commit, dim3(1), dim3(64), 0, st, s->gcnt, (int32_t)s->commit_pending);
    s->commit_pending = 0;
    LAUNCH_CHECK();
  }
  if (s->list_parity == 1) {
    // the last gather read parity 0's drift-margin flags, which still hold
    // the last tiled kernel's: cleared for the parity the next call starts with
    const int64_t m = s->nblocks;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, s->bmargin, 0u, m);
    LAUNCH_CHECK();
  }
  s->list_parity = 0;
  return AQ_OK;
}

// gather grid in list mode: enough wavefronts for the blocks of a compact
// scene, a loop over the list beyond that
static unsigned lblocks(const aq_sim *s) {
  const unsigned g = gblocks(s);
  return g < 2048u ? g : 2048u;
}

int launch_gather(aq_sim *s, bool tiled, bool op, uint32_t step, hipStream_t st) {
  const int32_t *ioff = tiled ? s->item_off : nullptr;
  const ColliderK K = collider_at(s, step);
  if (tiled) {
    if (s->commit_pending) {  // two list gathers with no tiled kernel between them
      hipLaunchKernelGGL(k_glist_commit, dim3(1), dim3(64), 0, st, s->gcnt, (int32_t)s->commit_pending);
      s->commit_pending = 0;
    }
    const int32_t fslot = 1 + s->list_parity;
    const uint32_t *bm_read = s->bmargin + (int64_t)s->list_parity * s->nblocks;
    uint32_t *bm_clear = s->bmargin + (int64_t)(1 - s->list_parity) * s->nblocks;
    dispatch_grid(s, [&](auto tag_) {
      TAG_TYPES(tag_)
      (void)MAT;
      (void)ST;
      (void)NW32;
      if (op)
        hipLaunchKernelGGL((k_gather_list<DIM, T, true>), dim3(lblocks(s)), dim3(256), 0, st, params_of<T>(s), K,
                           ioff, s->near, (const vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags,
                           (vec4_t<T> *)s->grid_v, s->act, s->glist, (const int32_t *)s->flist, s->gcnt,
                           s->listed, fslot, bm_read, bm_clear);
      else
        hipLaunchKernelGGL((k_gather_list<DIM, T, false>), dim3(lblocks(s)), dim3(256), 0, st, params_of<T>(s), K,
                           ioff, s->near, (const vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags,
                           (vec4_t<T> *)s->grid_v, s->act, s->glist, (const int32_t *)s->flist, s->gcnt,
                           s->listed, fslot, bm_read, bm_clear);
    });
    s->commit_pending = (uint32_t)fslot;
    s->list_parity ^= 1;
    LAUNCH_CHECK();
    return AQ_OK;
  }
  dispatch_grid(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    (void)MAT;
    (void)ST;
    (void)NW32;
    if (op)
      hipLaunchKernelGGL((k_gather<DIM, T, true>), dim3(gblocks(s)), dim3(256), 0, st, params_of<T>(s), K, s->nblocks,
                         ioff, s->near, (const vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags,
                         (vec4_t<T> *)s->grid_v, s->act);
    else
      hipLaunchKernelGGL((k_gather<DIM, T, false>), dim3(gblocks(s)), dim3(256), 0, st, params_of<T>(s), K, s->nblocks,
                         ioff, s->near, (const vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags,
                         (vec4_t<T> *)s->grid_v, s->act);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_p2g(aq_sim *s, const void *state, int64_t n, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  bool tiled;
  if ((rc = launch_p2g(s, state, n, &tiled, st))) return rc;
  if ((rc = launch_gather(s, tiled, false, 0u, st))) return rc;
  return settle_list(s, st);
}

extern "C" int aq_grid_op(aq_sim *s, uint32_t step, void *stream) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  hipStream_t st = (hipStream_t)stream;
  const ColliderK K = collider_at(s, step);
  dispatch_grid(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    (void)MAT;
    (void)ST;
    (void)NW32;
    hipLaunchKernelGGL((k_grid_update<DIM, T>), dim3(gblocks(s)), dim3(256), 0, st, params_of<T>(s), K, s->nblocks,
                       (vec4_t<T> *)s->grid_v, (const uint32_t *)s->act);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_g2p(aq_sim *s, void *state, int64_t n, const int32_t *ids, uint32_t step, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  return launch_g2p(s, state, n, ids, step, (hipStream_t)stream);
}

extern "C" int aq_sort(aq_sim *s, void *state, int64_t n, int32_t *ids, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  StateK sk = make_state(s, state, nullptr);
  CodecK cod = make_codec(s, 0u);
  const int32_t m = s->nblocks;
  hipLaunchKernelGGL(k_fill_u32, dim3((m + 1 + 255) / 256), dim3(256), 0, st, s->scount, 0u, (int64_t)m + 1);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)(((int64_t)m * 64 + 255) / 256)), dim3(256), 0, st, s->ccount, 0u,
                     (int64_t)m * 64);
  LAUNCH_CHECK();
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_sort_keys<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                       params_of<T>(s), n, s->skey, s->srank, s->scount, s->ccount);
  });
  hipLaunchKernelGGL(k_sort_rank, dim3(pblocks(n)), dim3(256), 0, st, s->skey, s->srank, s->ccount, n);
  LAUNCH_CHECK();
  const int32_t chunks = (m + 4095) / 4096;
  hipLaunchKernelGGL(k_scan_local, dim3(chunks), dim3(256), 0, st, s->scount, s->soff, s->item_off, s->spart, m);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, s->spart, chunks, s->soff, s->item_off, m);
  hipLaunchKernelGGL(k_scan_add, dim3(chunks), dim3(256), 0, st, s->soff, s->item_off, s->spart, m);
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
  // work items: up to kItem particles of one block each
  hipLaunchKernelGGL(k_items_fill, dim3((m + 255) / 256), dim3(256), 0, st, s->scount, s->soff, s->item_off,
                     s->items, m);
  // near flags; the gather list's counters (glist, both flist counters,
  // cursor) and the drift-margin masks of both parities cleared
  if (s->d.dim == 3)
    hipLaunchKernelGGL(k_near<3>, dim3((m + 255) / 256), dim3(256), 0, st, s->scount, (int32_t)s->pf.nb, s->near,
                       s->bmargin, s->gcnt, m);
  else
    hipLaunchKernelGGL(k_near<2>, dim3((m + 255) / 256), dim3(256), 0, st, s->scount, (int32_t)s->pf.nb, s->near,
                       s->bmargin, s->gcnt, m);
  s->commit_pending = 0;
  s->list_parity = 0;
  hipLaunchKernelGGL(k_glist_build, dim3((m + 255) / 256), dim3(256), 0, st, (const uint32_t *)s->near,
                     (const uint32_t *)s->act, s->listed, s->glist, s->gcnt, m);
  LAUNCH_CHECK();
  s->have_items = true;
  s->items_state = state;
  s->items_n = n;
  return AQ_OK;
}

extern "C" int aq_substep(aq_sim *s, void *state, int64_t n, int32_t *ids, uint32_t step, void *stream) {
  int rc;
  if ((rc = aq_sort(s, state, n, ids, stream))) return rc;
  if ((rc = aq_p2g(s, state, n, stream))) return rc;
  if ((rc = aq_grid_op(s, step, stream))) return rc;
  return aq_g2p(s, state, n, ids, step, stream);
}

// Event pairs around launches, summed per category when the call ends.
struct Timer {
  enum { FUSED = 0, GRID = 1, SORT = 2, ENDS = 3, NCAT = 4 };
  hipStream_t st;
  std::vector<hipEvent_t> a, b;
  std::vector<int> cat;
  std::vector<hipEvent_t> pool;  // created before the timed launches (event creation is host work
  size_t next = 0;               // that would otherwise sit between a start event and its kernel)
  int launches[NCAT] = {0, 0, 0, 0};
  bool ok = true;
  void reserve(size_t pairs) {
    pool.assign(2 * pairs, nullptr);
    for (auto &e : pool) ok = ok && hipEventCreate(&e) == hipSuccess;
  }
  hipEvent_t take() {
    if (next < pool.size()) {
      hipEvent_t e = pool[next];
      pool[next++] = nullptr;
      return e;
    }
    hipEvent_t e = nullptr;
    ok = ok && hipEventCreate(&e) == hipSuccess;
    return e;
  }
  void begin(int c) {
    hipEvent_t e0 = take(), e1 = take();
    ok = ok && e0 && e1 && hipEventRecord(e0, st) == hipSuccess;
    a.push_back(e0);
    b.push_back(e1);
    cat.push_back(c);
    launches[c]++;
  }
  void end() { ok = ok && hipEventRecord(b.back(), st) == hipSuccess; }
  // Per-category sums in ms; false when an event call failed.
  bool collect(float *ms) {
    ok = ok && hipStreamSynchronize(st) == hipSuccess;
    for (int c = 0; c < NCAT; ++c) ms[c] = 0.0f;
    for (size_t i = 0; i < a.size(); ++i) {
      float t = 0.0f;
      if (ok) ok = hipEventElapsedTime(&t, a[i], b[i]) == hipSuccess;
      ms[cat[i]] += t;
      if (a[i]) (void)hipEventDestroy(a[i]);
      if (b[i]) (void)hipEventDestroy(b[i]);
    }
    a.clear();
    b.clear();
    for (auto &e : pool)
      if (e) (void)hipEventDestroy(e);
    pool.clear();
    return ok;
  }
};

#define TIMED(TM, C, STMT)         \
  do {                             \
    if (TM) (TM)->begin(C);        \
    STMT;                          \
    if (TM) (TM)->end();           \
  } while (0)

// nsteps substeps: [sort] P2G, gather+grid update, then per substep the
// fused G2P->P2G and gather+grid update ([sort] every sort_every substeps),
// and a last G2P.
static int launch_limits(aq_sim *s, const void *state, int64_t n, float *mxmn, hipStream_t st);

// mxmn (optional): running per-component max / min of the state after every substep.
static int advance_impl(aq_sim *s, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                        int32_t *ids, hipStream_t st, Timer *tm, float *mxmn = nullptr) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (nsteps < 0) return fail(AQ_ERR_INVALID, "negative step count");
  if (nsteps == 0 || n == 0) return AQ_OK;
  if (sort_every > 0) {
    TIMED(tm, Timer::SORT, rc = aq_sort(s, state, n, ids, st));
    if (rc) return rc;
  }
  bool tiled;
  TIMED(tm, Timer::ENDS, rc = launch_p2g(s, state, n, &tiled, st));
  if (rc) return rc;
  TIMED(tm, Timer::GRID, rc = launch_gather(s, tiled, true, step0, st));
  if (rc) return rc;
  for (int t = 0; t + 1 < nsteps; ++t) {
    if (sort_every > 0 && t > 0 && t % sort_every == 0) {
      TIMED(tm, Timer::SORT, rc = aq_sort(s, state, n, ids, st));
      if (rc) return rc;
    }
    TIMED(tm, Timer::FUSED, rc = launch_g2p2g(s, state, n, ids, step0 + (uint32_t)t, &tiled, st));
    if (rc) return rc;
    if (mxmn && (rc = launch_limits(s, state, n, mxmn, st))) return rc;
    TIMED(tm, Timer::GRID, rc = launch_gather(s, tiled, true, step0 + (uint32_t)t + 1u, st));
    if (rc) return rc;
  }
  TIMED(tm, Timer::ENDS, rc = launch_g2p(s, state, n, ids, step0 + (uint32_t)(nsteps - 1), st));
  if (rc) return rc;
  if ((rc = settle_list(s, st))) return rc;
  if (mxmn) rc = launch_limits(s, state, n, mxmn, st);
  return rc;
}

extern "C" int aq_advance_limits(aq_sim *s, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                                 int32_t *ids, float *mxmn, void *stream) {
  if (!mxmn) return fail(AQ_ERR_INVALID, "null output");
  return advance_impl(s, state, n, nsteps, step0, sort_every, ids, (hipStream_t)stream, nullptr, mxmn);
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
  if (nsteps > 0) tm.reserve((size_t)nsteps * 2 + (size_t)(sort_every > 0 ? nsteps / sort_every + 1 : 0) + 4);
  int rc = advance_impl(s, state, n, nsteps, step0, sort_every, ids, st, &tm);
  const bool timed = tm.collect(ms);
  if (rc) return rc;
  if (!timed) return fail(AQ_ERR_HIP, "event timing failed");
  for (int c = 0; c < Timer::NCAT; ++c) launches[c] = tm.launches[c];
  // active grid nodes: nodes of the blocks of the last gather
  unsigned long long *cnt = (unsigned long long *)s->tmp_ids;  // free after the last sort
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)cnt, 0u, (int64_t)2);
  hipLaunchKernelGGL(k_count_active, dim3((s->nblocks + 255) / 256), dim3(256), 0, st, s->act, s->nblocks, cnt);
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
  return launch_limits(s, state, n, mxmn, (hipStream_t)stream);
}

static int launch_limits(aq_sim *s, const void *state, int64_t n, float *mxmn, hipStream_t st) {
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state, nullptr);
  CodecK cod = make_codec(s, 0u);
  const int32_t nparts = (int32_t)(pblocks(n) < (unsigned)kLimitBlocks ? pblocks(n) : (unsigned)kLimitBlocks);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_limits<DIM, MAT, ST, NW32, T>), dim3(nparts), dim3(256), 0, st, sk, cod, n, s->lim_part);
  });
  hipLaunchKernelGGL(k_limits_final, dim3(1), dim3(256), 0, st, s->lim_part, nparts, (int32_t)s->d.ncomp, mxmn);
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
  StateK sk = make_state(s, state, nullptr);
  CodecK cod = make_codec(s, 0u);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_energy<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                       params_of<T>(s), n, out);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

// ---- modelled quantization noise of the adjoint configuration ------------
// xi(step, c, p) = u - 1/2 with u the dithering noise of (seed, step,
// component c, particle id p): uniform in [-1/2, 1/2), 24-bit grid.  The id
// is ids[slot] when ids is given (a sorted state carries its particles'
// ids), else the slot.
struct SigmaK {
  double s[AQ_MAX_COMPS];
};

template <typename T>
__global__ void __launch_bounds__(256) k_noise_add(T *__restrict__ state, int64_t cap, int64_t n, int32_t nc,
                                                   const int32_t *__restrict__ ids, SigmaK sig, KeysK keys) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t id = ids ? (uint32_t)ids[p] : (uint32_t)p;
  for (int c = 0; c < nc; ++c) {
    if (sig.s[c] == 0.0) continue;
    const T xi = (T)(noise_u(id, keys.k[c]) - 0.5f);
    state[(int64_t)c * cap + p] += (T)sig.s[c] * xi;
  }
}

// out[c] += sum_p adj[c][p] xi(step, c, p), products and sums in double.
template <typename T>
__global__ void __launch_bounds__(256) k_noise_dot(const T *__restrict__ adj, int64_t cap, int64_t n, int32_t nc,
                                                   const int32_t *__restrict__ ids, KeysK keys,
                                                   double *__restrict__ out) {
  __shared__ double part[4][AQ_MAX_COMPS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t id = p < n ? (ids ? (uint32_t)ids[p] : (uint32_t)p) : 0u;
  for (int c = 0; c < nc; ++c) {
    double q = 0.0;
    if (p < n) q = (double)adj[(int64_t)c * cap + p] * (double)(noise_u(id, keys.k[c]) - 0.5f);
    for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o);
    if (lane == 0) part[wave][c] = q;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    const int c = threadIdx.x;
    atomicAdd(out + c, ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c]);
  }
}

extern "C" int aq_noise_add(aq_sim *s, void *state, int64_t n, const int32_t *ids, const double *sigma,
                            uint64_t seed, uint32_t step, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (s->d.storage == AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "noise is added to plain (f32/f64) storage");
  if (!sigma) return fail(AQ_ERR_INVALID, "null sigma");
  if (n == 0) return AQ_OK;
  SigmaK sig{};
  for (int c = 0; c < s->d.ncomp; ++c) {
    if (!(sigma[c] >= 0.0) || !std::isfinite(sigma[c])) return fail(AQ_ERR_INVALID, "sigma must be finite and >= 0");
    sig.s[c] = sigma[c];
  }
  const KeysK keys = make_keys(seed, step);
  hipStream_t st = (hipStream_t)stream;
  if (s->d.storage == AQ_STORE_F64)
    hipLaunchKernelGGL(k_noise_add<double>, dim3(pblocks(n)), dim3(256), 0, st, (double *)state, s->d.capacity, n,
                       s->d.ncomp, ids, sig, keys);
  else
    hipLaunchKernelGGL(k_noise_add<float>, dim3(pblocks(n)), dim3(256), 0, st, (float *)state, s->d.capacity, n,
                       s->d.ncomp, ids, sig, keys);
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_noise_dot(aq_sim *s, const void *adj, int64_t n, const int32_t *ids, uint64_t seed, uint32_t step,
                            double *out, void *stream) {
  int rc = check_n(s, adj, n);
  if (rc) return rc;
  if (s->d.storage == AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "the adjoint runs on plain (f32/f64) storage");
  if (s->has_collider || s->d.bound_clip)
    return fail(AQ_ERR_INVALID, "the adjoint of the SDF collider and of the bound clip is not implemented");
  if (!out) return fail(AQ_ERR_INVALID, "null output");
  if (n == 0) return AQ_OK;
  const KeysK keys = make_keys(seed, step);
  hipStream_t st = (hipStream_t)stream;
  if (s->d.storage == AQ_STORE_F64)
    hipLaunchKernelGGL(k_noise_dot<double>, dim3(pblocks(n)), dim3(256), 0, st, (const double *)adj, s->d.capacity,
                       n, s->d.ncomp, ids, keys, out);
  else
    hipLaunchKernelGGL(k_noise_dot<float>, dim3(pblocks(n)), dim3(256), 0, st, (const float *)adj, s->d.capacity, n,
                       s->d.ncomp, ids, keys, out);
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_sim_set_collider(aq_sim *s, const aq_collider *c) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  if (!c) {
    s->has_collider = false;
    return AQ_OK;
  }
  if (s->d.dim != 3) return fail(AQ_ERR_INVALID, "the SDF collider is 3D");
  if (!c->sdf || !c->normal) return fail(AQ_ERR_INVALID, "null collider volume");
  for (int a = 0; a < 3; ++a)
    if (c->size[a] < 1 || c->size[a] > 4096) return fail(AQ_ERR_INVALID, "collider volume size out of range");
  if (c->lag < 0 || c->lag > 1) return fail(AQ_ERR_INVALID, "collider lag must be 0 or 1");
  if (c->velocity_from_rotated != 0 && c->velocity_from_rotated != 1)
    return fail(AQ_ERR_INVALID, "velocity_from_rotated must be 0 or 1");
  if (!std::isfinite(c->omega) || !std::isfinite(c->height_rate)) return fail(AQ_ERR_INVALID, "non-finite collider motion");
  s->col = *c;
  s->has_collider = true;
  return AQ_OK;
}

extern "C" int aq_sim_count_wraps(aq_sim *s, unsigned long long *counts) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  if (counts && s->d.storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "only quantized records wrap");
  s->wraps = counts;
  return AQ_OK;
}

extern "C" int aq_sim_grid(aq_sim *s, void **grid, int64_t *nodes) {
  if (!s || !grid || !nodes) return fail(AQ_ERR_INVALID, "null argument");
  *grid = s->grid_v;
  *nodes = s->nodes;
  return AQ_OK;
}

// ---- codec self-check ------------------------------------------------------------
// dec_shared_enc (the encoder's fast decode of signed shared-exponent digits)
// against dec_shared, bit for bit, for every exponent field and digit
// pattern of one (digit bits, exponent bits) shape.
__global__ void __launch_bounds__(256) k_dec_shared_check(int nbits, int exp_bits, unsigned long long *bad) {
  const uint32_t n = 1u << (nbits + exp_bits);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CompK c = {};
  c.kind = K_FLOAT;
  c.nbits = (uint8_t)nbits;
  c.is_signed = 1;
  c.exp_bits = (uint8_t)exp_bits;
  c.shared = 1;
  c.conv = 127u - (1u << (exp_bits - 1)) + 1u;
  const uint32_t e = i >> nbits, q = i & ((1u << nbits) - 1u);
  if (__float_as_uint(dec_shared_enc(e, q, c)) != __float_as_uint(dec_shared(e, q, c))) atomicAdd(bad, 1ull);
}

extern "C" int aq_codec_selftest(unsigned long long *mismatches, void *stream) {
  if (!mismatches) return fail(AQ_ERR_INVALID, "null output");
  hipStream_t st = (hipStream_t)stream;
  for (int eb = 4; eb <= 8; ++eb)
    for (int nb = 2; nb <= 20; ++nb) {
      const uint32_t n = 1u << (nb + eb);
      hipLaunchKernelGGL(k_dec_shared_check, dim3((n + 255) / 256), dim3(256), 0, st, nb, eb, mismatches);
    }
  LAUNCH_CHECK();
  return AQ_OK;
}
