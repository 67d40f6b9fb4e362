// libaqmpm: groups of quantized simulations advanced together (aq_group_*).
#include <memory>

#include "aq_sim.hpp"
#include "aq_gather.hpp"

// The gather + grid update of a group of simulations in one launch: block
// numbers run simulation after simulation (bprefix[s] = blocks before
// simulation s), one block per wavefront.
template <int DIM>
__global__ void __launch_bounds__(256) k_gather_group(const GatherArgs<float> *__restrict__ descs,
                                                      const int32_t *__restrict__ bprefix, int32_t nsims) {
  const int32_t g = __builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6)));
  if (g >= bprefix[nsims]) return;
  int lo = 0, hi = nsims;  // bprefix[lo] <= g < bprefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bprefix[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  const GatherArgs<float> &A = descs[lo];
  ColliderK none{};  // group members have no collider
  gather_block<DIM, float, true>(A.P, none, g - bprefix[lo], A.item_off, A.near, A.tiles, A.grid_f, A.fflags, A.grid_v,
                                 A.act);
}

// ---- groups: many quantized simulations advanced together -----------------

// prefix[s] = items of the group's simulations before s (item_off[nblocks]
// of each, written by its last sort), prefix[nsims] = all of them.  One
// workgroup of 256.
__global__ void __launch_bounds__(256) k_group_prefix(const GatherArgs<float> *__restrict__ descs, int32_t nsims,
                                                      int32_t *__restrict__ prefix) {
  __shared__ int32_t wsum[4];
  __shared__ int32_t carry;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nsims; base += 256) {
    const int i = base + tid;
    const int32_t v = i < nsims ? descs[i].item_off[descs[i].nblocks] : 0;
    int32_t x = v;  // inclusive scan over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int32_t off = carry;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    if (i < nsims) prefix[i] = off + x - v;
    __syncthreads();  // every thread has read carry and wsum
    if (tid == 255) carry = off + x;
    __syncthreads();
  }
  if (tid == 0) prefix[nsims] = carry;
}

// Simulations of one fused launch (same dimension, material and word bucket).
struct GroupFused {
  int dim, mat, bucket;
  int32_t nsims;
  int64_t items_bound;       // sum over members of (n + kItem - 1) / kItem + nblocks
  TiledArgs<float> *descs;   // workspace: [nsims]
  GatherArgs<float> *gdescs; // workspace: [nsims], for the item prefix
  int32_t *prefix;           // workspace: [nsims + 1] items before each member
};

// Simulations of one gather launch (same dimension).
struct GroupGather {
  int dim;
  int32_t nsims;
  int32_t blocks;           // all members' grid blocks
  GatherArgs<float> *descs; // workspace: [nsims]
  int32_t *bprefix;         // workspace: [nsims + 1]
};

struct aq_group {
  std::vector<aq_sim *> sims;
  std::vector<void *> states;
  std::vector<int64_t> ns;
  std::vector<int32_t *> ids;
  std::vector<GroupFused> fused;
  std::vector<GroupGather> gathers;
};

static int64_t group_ws_total(int nsims) {
  // per simulation: one TiledArgs, two GatherArgs (fused class and gather
  // class) and two prefix entries; per array (at most 3 x 12 fused and
  // 2 x 2 gather arrays) the closing entry and the 256-byte alignment
  const int64_t per = (int64_t)sizeof(TiledArgs<float>) + 2 * (int64_t)sizeof(GatherArgs<float>) + 8;
  return align256((int64_t)nsims * per) + 48 * 256;
}

extern "C" int64_t aq_group_workspace_bytes(int nsims) {
  if (nsims < 1) return fail(AQ_ERR_INVALID, "a group needs a simulation");
  return group_ws_total(nsims);
}

static GatherArgs<float> gather_args(const aq_sim *s) {
  GatherArgs<float> a;
  a.P = s->pf;
  a.nblocks = s->nblocks;
  a.item_off = s->item_off;
  a.near = s->near;
  a.tiles = (const vec4_t<float> *)s->tiles;
  a.grid_f = (vec4_t<float> *)s->grid_f;
  a.fflags = s->fflags;
  a.grid_v = (vec4_t<float> *)s->grid_v;
  a.act = s->act;
  return a;
}

extern "C" int aq_group_create(aq_sim *const *sims, void *const *states, const int64_t *ns, int32_t *const *ids,
                               int nsims, void *workspace, int64_t workspace_bytes, aq_group **out) {
  if (!out) return fail(AQ_ERR_INVALID, "null output");
  *out = nullptr;
  if (!sims || !states || !ns) return fail(AQ_ERR_INVALID, "null argument");
  if (nsims < 1) return fail(AQ_ERR_INVALID, "a group needs a simulation");
  if (!workspace) return fail(AQ_ERR_INVALID, "null workspace");
  if (((uintptr_t)workspace & 255u) != 0u) return fail(AQ_ERR_INVALID, "workspace must be 256-byte aligned");
  if (workspace_bytes < group_ws_total(nsims)) return fail(AQ_ERR_CAPACITY, "group workspace too small");
  for (int i = 0; i < nsims; ++i) {
    const aq_sim *s = sims[i];
    if (!s) return fail(AQ_ERR_INVALID, "null simulation handle");
    if (s->d.storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "group members need quantized storage");
    if (s->d.compute_f64) return fail(AQ_ERR_INVALID, "group members compute in f32");
    if (s->has_collider) return fail(AQ_ERR_INVALID, "group members cannot have an SDF collider");
    if (ns[i] < 1 || ns[i] > s->d.capacity) return fail(AQ_ERR_CAPACITY, "particle count outside [1, capacity]");
    if (!states[i]) return fail(AQ_ERR_INVALID, "null state");
    for (int j = 0; j < i; ++j)
      if (sims[j] == s) return fail(AQ_ERR_INVALID, "a simulation appears twice in the group");
  }
  auto g = std::make_unique<aq_group>();
  g->sims.assign(sims, sims + nsims);
  g->states.assign(states, states + nsims);
  g->ns.assign(ns, ns + nsims);
  g->ids.assign(nsims, nullptr);
  if (ids) g->ids.assign(ids, ids + nsims);
  // classes, members in the caller's order
  std::vector<std::vector<int>> fmem, gmem;
  for (int i = 0; i < nsims; ++i) {
    const aq_sim *s = sims[i];
    const int b = nw32_bucket(s->d.nwords);
    int fc = -1;
    for (size_t c = 0; c < g->fused.size(); ++c)
      if (g->fused[c].dim == s->d.dim && g->fused[c].mat == s->d.material && g->fused[c].bucket == b) fc = (int)c;
    if (fc < 0) {
      GroupFused f{};
      f.dim = s->d.dim;
      f.mat = s->d.material;
      f.bucket = b;
      g->fused.push_back(f);
      fmem.emplace_back();
      fc = (int)g->fused.size() - 1;
    }
    fmem[fc].push_back(i);
    int gc = -1;
    for (size_t c = 0; c < g->gathers.size(); ++c)
      if (g->gathers[c].dim == s->d.dim) gc = (int)c;
    if (gc < 0) {
      GroupGather q{};
      q.dim = s->d.dim;
      g->gathers.push_back(q);
      gmem.emplace_back();
      gc = (int)g->gathers.size() - 1;
    }
    gmem[gc].push_back(i);
  }
  // workspace: descriptor arrays, then prefix arrays, each 256-aligned
  std::vector<char> host((size_t)workspace_bytes, 0);
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off = align256(off + bytes);
    return o;
  };
  std::vector<int64_t> at;  // offsets of the host copies to upload
  for (size_t c = 0; c < g->fused.size(); ++c) {
    GroupFused &f = g->fused[c];
    f.nsims = (int32_t)fmem[c].size();
    const int64_t od = take((int64_t)f.nsims * (int64_t)sizeof(TiledArgs<float>));
    const int64_t og = take((int64_t)f.nsims * (int64_t)sizeof(GatherArgs<float>));
    const int64_t op = take((int64_t)(f.nsims + 1) * 4);
    if (off > workspace_bytes) return fail(AQ_ERR_CAPACITY, "group workspace too small");
    f.descs = (TiledArgs<float> *)((char *)workspace + od);
    f.gdescs = (GatherArgs<float> *)((char *)workspace + og);
    f.prefix = (int32_t *)((char *)workspace + op);
    f.items_bound = 0;
    for (int k = 0; k < f.nsims; ++k) {
      const int i = fmem[c][k];
      aq_sim *s = g->sims[i];
      TiledArgs<float> a;
      a.s = make_state(s, g->states[i], g->ids[i]);
      a.cod = make_codec(s, 0u);
      // the seed's key; the kernel derives each substep's key from it
      a.cod.key_base = host_hash32((uint32_t)s->d.seed ^ host_hash32((uint32_t)(s->d.seed >> 32)));
      a.P = s->pf;
      a.n = g->ns[i];
      a.items = s->items;
      a.grid_v = (const vec4_t<float> *)s->grid_v;
      a.tiles = (vec4_t<float> *)s->tiles;
      a.grid_f = (vec4_t<float> *)s->grid_f;
      a.fflags = s->fflags;
      a.flist = nullptr;  // the group's gather visits every block
      a.fcnt = nullptr;
      a.bmargin = nullptr;  // and reads whole tiles
      a.item_off = s->item_off;
      std::memcpy(host.data() + od + (int64_t)k * (int64_t)sizeof(a), &a, sizeof(a));
      const GatherArgs<float> q = gather_args(s);
      std::memcpy(host.data() + og + (int64_t)k * (int64_t)sizeof(q), &q, sizeof(q));
      f.items_bound += (g->ns[i] + kItem - 1) / kItem + s->nblocks;
    }
  }
  for (size_t c = 0; c < g->gathers.size(); ++c) {
    GroupGather &q = g->gathers[c];
    q.nsims = (int32_t)gmem[c].size();
    const int64_t od = take((int64_t)q.nsims * (int64_t)sizeof(GatherArgs<float>));
    const int64_t op = take((int64_t)(q.nsims + 1) * 4);
    if (off > workspace_bytes) return fail(AQ_ERR_CAPACITY, "group workspace too small");
    q.descs = (GatherArgs<float> *)((char *)workspace + od);
    q.bprefix = (int32_t *)((char *)workspace + op);
    int64_t blocks = 0;
    for (int k = 0; k < q.nsims; ++k) {
      const aq_sim *s = g->sims[gmem[c][k]];
      const GatherArgs<float> a = gather_args(s);
      std::memcpy(host.data() + od + (int64_t)k * (int64_t)sizeof(a), &a, sizeof(a));
      const int32_t pb = (int32_t)blocks;
      std::memcpy(host.data() + op + (int64_t)k * 4, &pb, 4);
      blocks += s->nblocks;
    }
    if (blocks > (int64_t)INT32_MAX / 2) return fail(AQ_ERR_CAPACITY, "too many grid blocks in one group");
    q.blocks = (int32_t)blocks;
    std::memcpy(host.data() + op + (int64_t)q.nsims * 4, &q.blocks, 4);
  }
  HIP_TRY(hipMemcpy(workspace, host.data(), (size_t)off, hipMemcpyHostToDevice));
  *out = g.release();
  return AQ_OK;
}

extern "C" int aq_group_destroy(aq_group *g) {
  delete g;
  return AQ_OK;
}

template <class F>
static void dispatch_class(int dim, int mat, int bucket, F &&f) {
  if (dim == 3 && mat == AQ_FLUID) dispatch_quant<3, AQ_FLUID, float>(bucket, f);
  else if (dim == 3) dispatch_quant<3, AQ_ELASTIC, float>(bucket, f);
  else if (mat == AQ_FLUID) dispatch_quant<2, AQ_FLUID, float>(bucket, f);
  else dispatch_quant<2, AQ_ELASTIC, float>(bucket, f);
}

static int group_sort(aq_group *g, hipStream_t st) {
  for (size_t i = 0; i < g->sims.size(); ++i) {
    const int rc = aq_sort(g->sims[i], g->states[i], g->ns[i], g->ids[i], st);
    if (rc) return rc;
  }
  for (const GroupFused &f : g->fused)
    hipLaunchKernelGGL(k_group_prefix, dim3(1), dim3(256), 0, st, (const GatherArgs<float> *)f.gdescs, f.nsims,
                       f.prefix);
  LAUNCH_CHECK();
  return AQ_OK;
}

static int group_gather(aq_group *g, hipStream_t st) {
  for (const GroupGather &q : g->gathers) {
    const unsigned nb = (unsigned)((q.blocks + 3) / 4);
    if (q.dim == 3)
      hipLaunchKernelGGL(k_gather_group<3>, dim3(nb), dim3(256), 0, st, (const GatherArgs<float> *)q.descs,
                         (const int32_t *)q.bprefix, q.nsims);
    else
      hipLaunchKernelGGL(k_gather_group<2>, dim3(nb), dim3(256), 0, st, (const GatherArgs<float> *)q.descs,
                         (const int32_t *)q.bprefix, q.nsims);
  }
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_group_advance(aq_group *g, int nsteps, uint32_t step0, int sort_every, int workgroups,
                                void *stream) {
  if (!g) return fail(AQ_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(AQ_ERR_INVALID, "negative step count");
  if (sort_every < 1) return fail(AQ_ERR_INVALID, "a group sorts: sort_every must be positive");
  if (workgroups < 0 || workgroups > 8192 || workgroups % 8 != 0)
    return fail(AQ_ERR_INVALID, "workgroups must be a multiple of 8 in [0, 8192]");
  if (nsteps == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = group_sort(g, st))) return rc;
  for (size_t i = 0; i < g->sims.size(); ++i) {
    bool tiled;
    // whole tiles: the group's gather reads them so
    if ((rc = launch_p2g(g->sims[i], g->states[i], g->ns[i], &tiled, st, true))) return rc;
  }
  if ((rc = group_gather(g, st))) return rc;
  for (int t = 0; t + 1 < nsteps; ++t) {
    if (t > 0 && t % sort_every == 0 && (rc = group_sort(g, st))) return rc;
    for (const GroupFused &f : g->fused) {
      int64_t want = workgroups;
      if (want == 0) {
        want = f.items_bound + 64;
        if (want > 4096) want = 4096;
        want = (want + 7) / 8 * 8;
      }
      const uint32_t step = step0 + (uint32_t)t;
      dispatch_class(f.dim, f.mat, f.bucket, [&](auto tag_) {
        TAG_TYPES(tag_)
        (void)ST;
        hipLaunchKernelGGL((k_tiled_group<DIM, MAT, NW32>), dim3((unsigned)want), dim3(kItem), 0, st,
                           (const TiledArgs<float> *)f.descs, (const int32_t *)f.prefix, f.nsims, step);
      });
      LAUNCH_CHECK();
    }
    if ((rc = group_gather(g, st))) return rc;
  }
  for (size_t i = 0; i < g->sims.size(); ++i)
    if ((rc = launch_g2p(g->sims[i], g->states[i], g->ns[i], g->ids[i], step0 + (uint32_t)(nsteps - 1), st)))
      return rc;
  return AQ_OK;
}
