// Internal header of libaqmpm: the simulation handle and the dispatch and
// launch helpers shared by the library's translation units (aq_mpm.hip,
// aq_particles.hip, aq_fused.hip, aq_group.hip, aq_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "aq_mpm.h"
#include "aq_mpm_kernels.hpp"
#include "aq_tiled.hpp"

using namespace aq;

// The calling thread's error message (aq_last_error); returns `code`.
__attribute__((visibility("hidden"))) int aq_fail(int code, const char *msg);
static inline int fail(int code, const std::string &msg) { return aq_fail(code, msg.c_str()); }

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(AQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

struct KeysK {
  uint32_t k[AQ_MAX_COMPS];
};

// Keys of one (seed, step): key_j = h(j ^ key_base), key_base = h(step ^
// h(seed_lo ^ h(seed_hi))).  Dithering: component c uses key_{c >> 1} (one
// hash per component pair, dither_u); modelled noise: component c uses
// key_c (noise_u).
static inline uint32_t host_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
static inline uint32_t key_base(uint64_t seed, uint32_t step) {
  const uint32_t k = host_hash32((uint32_t)seed ^ host_hash32((uint32_t)(seed >> 32)));
  return host_hash32(step ^ k);
}
static inline KeysK make_keys(uint64_t seed, uint32_t step) {
  KeysK K;
  const uint32_t k = key_base(seed, step);
  for (int c = 0; c < AQ_MAX_COMPS; ++c) K.k[c] = host_hash32((uint32_t)c ^ k);
  return K;
}

static inline int nw32_bucket(int nw) { return nw <= 2 ? 4 : nw <= 4 ? 8 : nw <= 8 ? 16 : 32; }

// The SDF collider during one substep (aq_collider at that substep).
struct ColliderK {
  const float *sdf;
  const float *normal;
  int32_t size[3], offset[3], center[3];
  int32_t on;        // a collider is set
  int32_t yshift;    // int(height / dx + 0.5)
  int32_t from_rot;  // collider velocity from the rotated index
  double cos_t, sin_t, omega;
};

// Grids: grid_v holds the gathered P2G result ((m v, m) after aq_p2g,
// velocities after the grid update) on the blocks flagged in `act`, zeros
// elsewhere; grid_f is the float-atomic accumulation grid of untiled
// scatters (zero outside the blocks flagged in fflags); `tiles` holds one
// P2G tile per work item.
struct aq_sim {
  aq_sim_desc d;
  LayoutK lay;
  CompK *lay_dev;  // lay.c in the workspace, read by the MPM kernels
  double *lay_f64; // f64 compute: {1 / scale, scale, offset, 0} per component (workspace)
  ParamsT<float> pf;
  ParamsT<double> pd;
  int64_t nodes;
  int32_t nblocks;
  int32_t node_bytes;
  void *grid_v, *grid_f, *tiles;
  uint32_t *fflags, *act;
  // gather list (tiled path): gcnt[0] blocks in glist (near a particle or
  // active at the last sort, plus blocks first reached since); gcnt[1] and
  // gcnt[2] count the blocks in flist that untiled scatters flagged, for
  // the next gather of parity 0 and 1 (tiled kernels append into the next
  // gather's counter while the last gather's counter waits to be cleared);
  // gcnt[3] the append cursor of glist; listed[b]: b is in glist.
  // After a list gather, commit_pending is the counter it consumed: the next
  // tiled kernel (or k_glist_commit) sets gcnt[0] = gcnt[3] and clears it.
  uint32_t *gcnt, *listed;
  // per grid block, two arrays (parity of the gather that reads them): bit i
  // set when the block's item i wrote its tile's drift margin (bit 31: items
  // 31 and later, which always write it)
  uint32_t *bmargin;
  int32_t list_parity = 0;     // parity of the next list gather
  uint32_t commit_pending = 0; // counter slot to clear at the next commit (0: none)
  int32_t *glist, *flist;
  uint32_t *skey, *srank, *scount, *ccount, *soff, *spart;
  void *tmp_state;
  int32_t *tmp_ids;
  int4 *items;
  int32_t *item_off;  // [nblocks + 1]; item_off[nblocks] = item count
  uint32_t *near;     // [nblocks]: particles within one block (k_near, at each sort)
  float *lim_part;    // [kLimitBlocks][2 ncomp]: partials of k_limits
  int32_t max_items;
  bool have_items;    // items describe the current order of `items_state`
  const void *items_state;
  int64_t items_n;
  bool dither_any;
  bool stream_order;  // the encoder's field order has non-decreasing offsets
  int64_t ws_bytes;
  hipModule_t jit_mod = nullptr;      // fused kernel specialised to the layout (aq_sim_specialize)
  hipFunction_t jit_fn = nullptr;      // fused G2P -> P2G
  hipFunction_t jit_fn_p2g = nullptr;  // P2G only (the first substep of a run)
  hipFunction_t jit_fn_g2p = nullptr;  // G2P only (the last substep of a run)
  unsigned long long *wraps = nullptr;  // range-violation counters (aq_sim_count_wraps)
  bool has_collider = false;            // aq_sim_set_collider
  aq_collider col{};
};

static inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

// Dispatch over (dim, material, storage, record bucket): calls f(Tag<...>{}).
template <int D_, int M_, int S_, int W_, class T_> struct Tag {
  static constexpr int DIM = D_, MAT = M_, ST = S_, NW32 = W_;
  using T = T_;
};

template <int D_, int M_, class T_, class F>
static inline void dispatch_quant(int bucket, F &&f) {
  if (bucket == 4) f(Tag<D_, M_, AQ_STORE_QUANT, 4, T_>{});
  else if (bucket == 8) f(Tag<D_, M_, AQ_STORE_QUANT, 8, T_>{});
  else if (bucket == 16) f(Tag<D_, M_, AQ_STORE_QUANT, 16, T_>{});
  else f(Tag<D_, M_, AQ_STORE_QUANT, 32, T_>{});
}

// Whether the simulation computes in double: plain f64 storage, or
// quantized records with compute_f64 (2D only).
static inline bool f64_compute(const aq_sim_desc &d) {
  return d.storage == AQ_STORE_F64 || (d.storage == AQ_STORE_QUANT && d.compute_f64);
}

template <class F>
static inline void dispatch(const aq_sim *s, F &&f) {
  const aq_sim_desc &d = s->d;
  if (d.storage == AQ_STORE_QUANT && d.compute_f64) {
    const int b = nw32_bucket(d.nwords);
    if (d.material == AQ_FLUID) dispatch_quant<2, AQ_FLUID, double>(b, f);
    else dispatch_quant<2, AQ_ELASTIC, double>(b, f);
  } else if (d.storage == AQ_STORE_QUANT) {
    const int b = nw32_bucket(d.nwords);
    if (d.dim == 3 && d.material == AQ_FLUID) dispatch_quant<3, AQ_FLUID, float>(b, f);
    else if (d.dim == 3) dispatch_quant<3, AQ_ELASTIC, float>(b, f);
    else if (d.material == AQ_FLUID) dispatch_quant<2, AQ_FLUID, float>(b, f);
    else dispatch_quant<2, AQ_ELASTIC, float>(b, f);
  } else if (d.storage == AQ_STORE_F32) {
    if (d.dim == 3 && d.material == AQ_FLUID) f(Tag<3, AQ_FLUID, AQ_STORE_F32, 4, float>{});
    else if (d.dim == 3) f(Tag<3, AQ_ELASTIC, AQ_STORE_F32, 4, float>{});
    else if (d.material == AQ_FLUID) f(Tag<2, AQ_FLUID, AQ_STORE_F32, 4, float>{});
    else f(Tag<2, AQ_ELASTIC, AQ_STORE_F32, 4, float>{});
  } else {
    if (d.dim == 3 && d.material == AQ_FLUID) f(Tag<3, AQ_FLUID, AQ_STORE_F64, 4, double>{});
    else if (d.dim == 3) f(Tag<3, AQ_ELASTIC, AQ_STORE_F64, 4, double>{});
    else if (d.material == AQ_FLUID) f(Tag<2, AQ_FLUID, AQ_STORE_F64, 4, double>{});
    else f(Tag<2, AQ_ELASTIC, AQ_STORE_F64, 4, double>{});
  }
}

// Dispatch over (dim, compute type) for the grid kernels.
template <class F>
static inline void dispatch_grid(const aq_sim *s, F &&f) {
  if (f64_compute(s->d)) {
    if (s->d.dim == 3) f(Tag<3, 0, 0, 0, double>{});
    else f(Tag<2, 0, 0, 0, double>{});
  } else {
    if (s->d.dim == 3) f(Tag<3, 0, 0, 0, float>{});
    else f(Tag<2, 0, 0, 0, float>{});
  }
}

#define TAG_TYPES(tag)                      \
  using TG = decltype(tag);                 \
  constexpr int DIM = TG::DIM;              \
  constexpr int MAT = TG::MAT;              \
  constexpr int ST = TG::ST;                \
  constexpr int NW32 = TG::NW32;            \
  using T = typename TG::T;

template <class T> const ParamsT<T> &params_of(const aq_sim *s);
template <> inline const ParamsT<float> &params_of<float>(const aq_sim *s) { return s->pf; }
template <> inline const ParamsT<double> &params_of<double>(const aq_sim *s) { return s->pd; }

static inline unsigned pblocks(int64_t n) { return (unsigned)((n + 255) / 256); }
static inline unsigned gblocks(const aq_sim *s) { return (unsigned)((s->nblocks + 3) / 4); }
// persistent grid of the tiled kernels: a multiple of 8 (one share per XCD)
static inline unsigned tblocks(const aq_sim *s, int64_t n) {
  int64_t want = (n + kItem - 1) / kItem + 64;
  if (want > 4096) want = 4096;
  return (unsigned)((want + 7) / 8 * 8);
}

static inline int check_n(const aq_sim *s, const void *state, int64_t n) {
  if (!s) return fail(AQ_ERR_INVALID, "null handle");
  if (n < 0 || n > s->d.capacity) return fail(AQ_ERR_CAPACITY, "particle count exceeds capacity");
  if (n && !state) return fail(AQ_ERR_INVALID, "null state");
  return AQ_OK;
}

// Control word of a tiled kernel of one simulation: bits 0-1 the counter
// slot its untiled-scatter appends go to (the next gather's), bits 2-3 the
// slot to clear while committing the last gather's appends (0: none), which
// the kernel's first thread does before any append.
static inline uint32_t take_list_ctl(aq_sim *s) {
  const uint32_t ctl = (uint32_t)(1 + s->list_parity) | (s->commit_pending << 2);
  s->commit_pending = 0;
  return ctl;
}

static inline bool items_valid(const aq_sim *s, const void *state, int64_t n) {
  return s->have_items && s->items_state == state && s->items_n == n;
}

// aq_mpm.hip
StateK make_state(const aq_sim *s, const void *state, const int32_t *ids);
CodecK make_codec(const aq_sim *s, uint32_t step);
int launch_gather(aq_sim *s, bool tiled, bool op, uint32_t step, hipStream_t st);
// the collider of substep `step` (off when none is set or in 2D)
ColliderK collider_at(const aq_sim *s, uint32_t step);
// Ends a call's use of the gather list: commits the last list gather if no
// tiled kernel did, then both flist counters are zero and the next gather
// takes parity 0 (every call leaves this state, so that a captured call
// replays on the state it was recorded from).
int settle_list(aq_sim *s, hipStream_t st);
// aq_particles.hip
int launch_p2g(aq_sim *s, const void *state, int64_t n, bool *tiled, hipStream_t st, bool whole = false);
int launch_g2p(aq_sim *s, void *state, int64_t n, const int32_t *ids, uint32_t step, hipStream_t st);
// aq_fused.hip
int launch_g2p2g(aq_sim *s, void *state, int64_t n, const int32_t *ids, uint32_t step, bool *tiled,
                 hipStream_t st);
