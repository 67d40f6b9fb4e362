// Gather of the P2G tiles into the grid, the grid update and the SDF
// collider (device code of libaqmpm, included by aq_mpm.hip and
// aq_group.hip).
#pragma once
#include "aq_sim.hpp"

// The collider at node I (rot_scene, mpm_solver.py:1524-1545, and sdf,
// :1447-1452): whether I maps to an inside voxel, and then the voxel's
// normal rotated by theta (n) and the collider velocity omega e_y x r (cv).
template <class T>
__device__ __forceinline__ bool collider_probe(const ParamsT<T> &P, const ColliderK &K, const int (&I)[3],
                                               T (&n)[3], T (&cv)[3]) {
  int w[3] = {I[0], I[1], I[2]};
  if (w[0] >= 2 * K.center[0]) w[0] -= 2 * K.center[0];
  if (w[2] >= 2 * K.center[2]) w[2] -= 2 * K.center[2];
  const T c = (T)K.cos_t, sn = (T)K.sin_t;
  const T o0 = (T)(w[0] - K.center[0]), o2 = (T)(w[2] - K.center[2]);
  // the offset rotated by -theta; indices truncated as ti.cast does, then
  // clamped to the walls
  const T xr = c * o0 - sn * o2 + (T)K.center[0] + (T)0.5;
  const T zr = sn * o0 + c * o2 + (T)K.center[2] + (T)0.5;
  int id[3] = {(int)xr, I[1] - K.yshift, (int)zr};
#pragma unroll
  for (int a = 0; a < 3; ++a) id[a] = id[a] > P.ub[a] ? P.ub[a] : (id[a] < P.lb[a] ? P.lb[a] : id[a]);
  int q[3];
  bool vol = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q[a] = id[a] - K.offset[a];
    vol = vol && q[a] >= 0 && q[a] < K.size[a];
  }
  if (!vol) return false;  // outside the volume: the stored value is 0, not inside
  const int64_t e = ((int64_t)q[0] * K.size[1] + q[1]) * K.size[2] + q[2];
  if (!(K.sdf[e] < -1e-5f)) return false;
  const T g0 = (T)K.normal[3 * e], g1 = (T)K.normal[3 * e + 1], g2 = (T)K.normal[3 * e + 2];
  n[0] = c * g0 + sn * g2;
  n[1] = g1;
  n[2] = -sn * g0 + c * g2;
  const int p0 = K.from_rot ? id[0] : I[0], p2 = K.from_rot ? id[2] : I[2];
  const T om = (T)K.omega;
  const T r0 = (T)(p0 - K.center[0]) * P.dx, r2 = (T)(p2 - K.center[2]) * P.dx;
  cv[0] = om * r2;  // omega e_y x (r0, 0, r2)
  cv[1] = (T)0;
  cv[2] = -(om * r0);
  return true;
}

// sdf_collision(_no_arg) of node I (mpm_solver.py:1459-1521) on the updated
// velocity v: v = cv + (v_rel - min(v_rel . n, 0) n), v_rel = v - cv.
template <class T>
__device__ __forceinline__ void collide_node(const ParamsT<T> &P, const ColliderK &K, const int (&I)[3],
                                             T (&v)[3]) {
  T n[3], cv[3];
  if (!collider_probe<T>(P, K, I, n, cv)) return;
  T vr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) vr[a] = v[a] - cv[a];
  const T dn = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
  const T m = tmin<T>(dn, (T)0);
#pragma unroll
  for (int a = 0; a < 3; ++a) v[a] = cv[a] + (vr[a] - m * n[a]);
}

// Grid update (mpm_solver.py:633-665) of node I with accumulated (m v, m),
// then the SDF collider (3D) when one is set.
template <int DIM, class T>
__device__ __forceinline__ vec4_t<T> grid_op_node(const ParamsT<T> &P, const ColliderK &K, const int (&I)[3],
                                                  vec4_t<T> nd) {
  const T inv_m = (T)1 / (nd.w + P.eps_m);
  T v[3] = {inv_m * nd.x, inv_m * nd.y, inv_m * nd.z};
  v[1] = v[1] - P.g_dt;
  if (P.boundary == AQ_STICKY) {
    bool wall = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) wall = wall || I[d] < P.lb[d] || I[d] > P.ub[d];
    if (wall) v[0] = v[1] = v[2] = (T)0;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      if (I[d] < P.lb[d] && v[d] < (T)0) v[d] = (T)0;
      if (I[d] > P.ub[d] && v[d] > (T)0) v[d] = (T)0;
    }
  }
  if constexpr (DIM == 3)
    if (K.on) collide_node<T>(P, K, I, v);
  nd.x = v[0];
  nd.y = v[1];
  nd.z = DIM == 3 ? v[2] : (T)0;
  return nd;
}

// Local coordinates of lane `lane` in a 64-node block.
template <int DIM> __device__ __forceinline__ void lane_coords(int lane, int (&l)[3]) {
  if constexpr (DIM == 3) {
    l[0] = lane >> 4;
    l[1] = (lane >> 2) & 3;
    l[2] = lane & 3;
  } else {
    l[0] = lane >> 3;
    l[1] = lane & 7;
    l[2] = 0;
  }
}

// Gather of P2G results into grid_v, one 64-node block per wavefront.  A
// node is covered by the tiles of 2^DIM neighbouring blocks (per axis its own
// block and the one on the side of the block edge it is nearest); each lane
// reads the item ranges of its 2^DIM blocks and the first two items' tile
// values all at once (independent loads), then any further items, and sums
// in a fixed order (neighbour, then item); grid_f is added for blocks flagged by untiled scatters (and zeroed after
// use).  OP applies the grid update; without it grid_v holds (m v, m).
// Blocks with no contribution are marked inactive in `act` (and zeroed if
// they were active).
// Everything the gather needs of one simulation.
template <class T> struct GatherArgs {
  ParamsT<T> P;
  int32_t nblocks;
  const int32_t *item_off;  // null: no tiles (untiled P2G)
  const uint32_t *near;
  const vec4_t<T> *tiles;
  vec4_t<T> *grid_f;
  uint32_t *fflags;
  vec4_t<T> *grid_v;
  uint32_t *act;
};

// Gather of block b (uniform over the wavefront) of one simulation.
// bm_read (null: every tile whole): per block, bit i set when its item i
// wrote its tile's drift margin (bit 31: items 31 and later, which always
// do); the margin nodes of the other items were not written and are not
// read.  bm_clear (null: none): the flags the next tiled kernel sets,
// cleared here for block b (every block with items is gathered).
// SPEC: the neighbours' item ranges are loaded with the block's flags
// instead of after them (one round trip less for a block that has items
// near it, wasted loads for one that has not: for gathers over lists of
// blocks near particles).
template <int DIM, class T, bool OP, bool SPEC = false>
__device__ __forceinline__ void gather_block(const ParamsT<T> &P, const ColliderK &K, int32_t b,
                                             const int32_t *__restrict__ item_off,
                                             const uint32_t *__restrict__ near, const vec4_t<T> *__restrict__ tiles,
                                             vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags,
                                             vec4_t<T> *__restrict__ grid_v, uint32_t *__restrict__ act,
                                             const uint32_t *__restrict__ bm_read = nullptr,
                                             uint32_t *__restrict__ bm_clear = nullptr) {
  using TL = Tile<DIM>;
  constexpr int E = TL::E, TN = TL::TN, NN = 1 << DIM;
  const int lane = threadIdx.x & 63;
  if (bm_clear && lane == 0) bm_clear[b] = 0u;
  const bool fall = fflags[b] != 0u;
  // a block with no items in its neighbourhood and no untiled scatter has
  // nothing to gather: skip the item ranges of its neighbours (with SPEC
  // their empty ranges lead to the same exit below)
  if (!SPEC && item_off && !near[b] && !fall) {
    if (act[b]) {
      vec4_t<T> z;
      z.x = z.y = z.z = z.w = (T)0;
      grid_v[(int64_t)b * 64 + lane] = z;
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) act[b] = 0u;
    }
    return;
  }
  const int nb = P.nb;
  int bc[3], loc[3];
  block_coords<DIM>(b, nb, bc);
  lane_coords<DIM>(lane, loc);
  int32_t r0[NN], r1[NN], q[NN];
  bool mg[NN];  // the node is in the drift margin of direction k's tiles
  uint32_t bf[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    int d[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    bool ok = item_off != nullptr;
    mg[k] = false;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      d[a] = ((k >> a) & 1) ? (loc[a] == E - 1 ? 1 : -1) : 0;
      const int c = bc[a] + d[a];
      t[a] = loc[a] + 1 - d[a] * E;
      ok = ok && c >= 0 && c < nb && t[a] < TL::TE;  // 2D: only 3 nodes reach past the edge
      mg[k] = mg[k] || t[a] == 0 || t[a] == TL::TE - 1;
    }
    q[k] = TL::store_pos(t[0], t[1], t[2]);
    r0[k] = r1[k] = 0;
    bf[k] = 0xFFFFFFFFu;
    if (ok) {
      const int32_t nbid = DIM == 3 ? ((bc[0] + d[0]) * nb + (bc[1] + d[1])) * nb + (bc[2] + d[2])
                                    : (bc[0] + d[0]) * nb + (bc[1] + d[1]);
      r0[k] = item_off[nbid];
      r1[k] = item_off[nbid + 1];
      if (bm_read && mg[k]) bf[k] = bm_read[nbid];
    }
  }
  // items whose margin was not written contribute nothing there
#pragma unroll
  for (int k = 0; k < NN; ++k)
    if (!mg[k]) bf[k] = 0xFFFFFFFFu;
  bool mine = false;
#pragma unroll
  for (int k = 0; k < NN; ++k) mine = mine || r1[k] > r0[k];
  const bool any = fall || __ballot(mine) != 0ull;
  const int64_t node = (int64_t)b * 64 + lane;
  if (!any) {
    if (act[b]) {
      vec4_t<T> z;
      z.x = z.y = z.z = z.w = (T)0;
      grid_v[node] = z;
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) act[b] = 0u;
    }
    return;
  }
  vec4_t<T> c0[NN], c1[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    c0[k].x = c0[k].y = c0[k].z = c0[k].w = (T)0;
    c1[k] = c0[k];
    if (r1[k] > r0[k] && (bf[k] & 1u)) c0[k] = tiles[(int64_t)r0[k] * TN + q[k]];
    if (r1[k] > r0[k] + 1 && (bf[k] & 2u)) c1[k] = tiles[(int64_t)(r0[k] + 1) * TN + q[k]];
  }
  vec4_t<T> sum;
  sum.x = sum.y = sum.z = sum.w = (T)0;
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    sum.x += c0[k].x;
    sum.y += c0[k].y;
    sum.z += c0[k].z;
    sum.w += c0[k].w;
    sum.x += c1[k].x;
    sum.y += c1[k].y;
    sum.z += c1[k].z;
    sum.w += c1[k].w;
    for (int32_t it = r0[k] + 2; it < r1[k]; ++it) {
      const int ib = it - r0[k];
      if (!((bf[k] >> (ib < 31 ? ib : 31)) & 1u)) continue;
      const vec4_t<T> c = tiles[(int64_t)it * TN + q[k]];
      sum.x += c.x;
      sum.y += c.y;
      sum.z += c.z;
      sum.w += c.w;
    }
  }
  if (fall) {
    const vec4_t<T> c = grid_f[node];
    sum.x += c.x;
    sum.y += c.y;
    sum.z += c.z;
    sum.w += c.w;
    vec4_t<T> z;
    z.x = z.y = z.z = z.w = (T)0;
    grid_f[node] = z;
  }
  if constexpr (OP) {
    int I[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) I[a] = bc[a] * E + loc[a];
    sum = grid_op_node<DIM, T>(P, K, I, sum);
  }
  grid_v[node] = sum;
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    act[b] = 1u;
    if (fall) fflags[b] = 0u;
  }
}
