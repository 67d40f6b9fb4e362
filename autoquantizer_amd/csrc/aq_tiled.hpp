// LDS-tiled fused G2P/P2G kernel over work items (gfx950), shared by the
// library build and the run-time specialisation of the kernel to one record
// layout (aq_sim_specialize).
#pragma once

#include "aq_mpm_kernels.hpp"

namespace aq {

// ---- LDS-tiled kernels over work items ---------------------------------------
// A work item is up to kItem particles whose stencil base was in one grid block
// at the last aq_sort: int4 (block, first particle, count, 0).  The item's
// tile covers the block, the 2-node stencil reach and a 1-cell drift margin
// on each side: TE = E + 4 nodes per axis from block_origin - 1 (LDS layout
// in Tile; in the tile buffer, TE^DIM nodes grouped by the grid block they
// belong to, Tile::store_pos).
//
// G2P reads grid velocities staged in LDS.  P2G accumulates in LDS in fixed
// point (integer LDS atomics, exact and so independent of the order): per
// item and component the contributions are scaled by 2^k, with
// 2^(BITS-k) above the item's summed contribution bound, so the integer sums
// cannot overflow and are exact sums of the terms rounded to 2^-BITS of that
// bound (BITS = 30 in int32 for f32 compute, 62 in int64 for f64).  The
// finished tile goes to the item's slot of the tile buffer with plain stores;
// the gather kernel sums the tiles covering each node in a fixed order.  A
// particle whose stencil leaves the tile (it drifted more than a cell since
// the sort), or an item with a non-finite bound, reads grid_v and scatters
// into grid_f.
// Particles per work item = threads of a tiled workgroup (4 wavefronts).
// 512 (8 wavefronts) halves the tiles and made the gather 7.5 µs faster on
// the 3D fluid benchmark, but the fused kernel 6-18 µs slower (fewer
// workgroups between barriers).
constexpr int kItem = 256;

template <int DIM> struct Tile {
  static constexpr int E = BlockShape<DIM>::E, LG = BlockShape<DIM>::LOG;
  static constexpr int TE = E + 4;                                // nodes per axis
  static constexpr int TN = DIM == 3 ? TE * TE * TE : TE * TE;    // nodes
  static constexpr int STRIDE = TN;                               // nodes per item in the tile buffer
  // LDS slot of tile node (x, y, z): x SX + y SY + z.  SX = 112 = 48 mod 64,
  // SY = 12 (3D) / SX = 72 = 8 mod 64 (2D): the nodes of any 4x4x4 (8x8)
  // window get 64 distinct banks, and stencil offsets are immediate offsets.
  static constexpr int SY = DIM == 3 ? 12 : 1;
  static constexpr int SX = DIM == 3 ? 112 : 72;
  static constexpr int TS = TE * SX;                              // LDS slots (< 1024)
  __host__ __device__ static constexpr int slot(int x, int y, int z) {
    return DIM == 3 ? x * SX + y * SY + z : x * SX + y;
  }
  __device__ static void node_coords(int nd, int (&t)[3]) {
    if constexpr (DIM == 3) {
      t[0] = nd / (TE * TE);
      t[1] = (nd / TE) % TE;
      t[2] = nd % TE;
    } else {
      t[0] = nd / TE;
      t[1] = nd % TE;
      t[2] = 0;
    }
  }
  // Position of tile node (x, y, z) in the tile buffer.  Per axis the tile
  // covers the last node of the previous block (t = 0), the E nodes of its
  // own block (1..E) and the first 3 of the next (E+1..E+3); the buffer
  // groups the nodes by the block they belong to (3^DIM groups, in group
  // order) and stores each group row-major, which is the gathering
  // wavefront's lane order: a wavefront reads each covering tile as one
  // contiguous run.
  __host__ __device__ static constexpr int store_pos(int x, int y, int z) {
    int g[3] = {0, 0, 0}, u[3] = {0, 0, 0}, sz[3] = {0, 0, 0};
    const int t[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g[a] = t[a] == 0 ? 0 : (t[a] <= E ? 1 : 2);
      u[a] = t[a] == 0 ? 0 : (t[a] <= E ? t[a] - 1 : t[a] - E - 1);
      sz[a] = g[a] == 0 ? 1 : (g[a] == 1 ? E : 3);
    }
    // nodes of the groups before (g0, g1, g2): the slab below g0, the rows
    // of slab g0 below g1, the groups of row (g0, g1) below g2
    const int cum[3] = {0, 1, E + 1};  // nodes per axis before group g
    if constexpr (DIM == 3) {
      const int before = cum[g[0]] * TE * TE + sz[0] * cum[g[1]] * TE + sz[0] * sz[1] * cum[g[2]];
      return before + (u[0] * sz[1] + u[1]) * sz[2] + u[2];
    } else {
      const int before = cum[g[0]] * TE + sz[0] * cum[g[1]];
      return before + u[0] * sz[1] + u[1];
    }
  }
};

// LDS slot of the node at each position of the tile buffer (the inverse of
// Tile::store_pos), built at compile time: the write-out walks the buffer
// in order, so each wavefront stores a contiguous run.
// Bit 15 marks the drift margin: nodes with t = 0 or TE - 1 on some axis,
// which only particles that left their sort cell reach.
constexpr uint16_t kMarginBit = 0x8000u;
template <int DIM> struct TileOrder {
  uint16_t slot[Tile<DIM>::TN];
};
template <int DIM> constexpr TileOrder<DIM> make_tile_order() {
  using TL = Tile<DIM>;
  TileOrder<DIM> o{};
  for (int x = 0; x < TL::TE; ++x)
    for (int y = 0; y < TL::TE; ++y)
      for (int z = 0; z < (DIM == 3 ? TL::TE : 1); ++z) {
        const bool margin = x == 0 || x == TL::TE - 1 || y == 0 || y == TL::TE - 1 ||
                            (DIM == 3 && (z == 0 || z == TL::TE - 1));
        o.slot[TL::store_pos(x, y, z)] = (uint16_t)(TL::slot(x, y, z) | (margin ? kMarginBit : 0u));
      }
  return o;
}
__constant__ constexpr TileOrder<3> kTileOrder3 = make_tile_order<3>();
__constant__ constexpr TileOrder<2> kTileOrder2 = make_tile_order<2>();
template <int DIM> __device__ __forceinline__ int tile_slot_at(int pos) {
  if constexpr (DIM == 3)
    return kTileOrder3.slot[pos];
  else
    return kTileOrder2.slot[pos];
}

template <int DIM>
__device__ __forceinline__ void block_coords(int32_t b, int nb, int (&bc)[3]) {
  if constexpr (DIM == 3) {
    bc[0] = b / (nb * nb);
    bc[1] = (b / nb) % nb;
    bc[2] = b % nb;
  } else {
    bc[0] = b / nb;
    bc[1] = b % nb;
    bc[2] = 0;
  }
}

template <int DIM>
__device__ __forceinline__ int64_t node_index(int i, int j, int k, int nb) {
  constexpr int LG = BlockShape<DIM>::LOG, E = BlockShape<DIM>::E;
  if constexpr (DIM == 3) {
    const int64_t blk = ((int64_t)(i >> LG) * nb + (j >> LG)) * nb + (k >> LG);
    return blk * 64 + (((i & (E - 1)) << 4) | ((j & (E - 1)) << 2) | (k & (E - 1)));
  } else {
    const int64_t blk = (int64_t)(i >> LG) * nb + (j >> LG);
    return blk * 64 + (((i & (E - 1)) << 3) | (j & (E - 1)));
  }
}

template <class T> __device__ __forceinline__ T tldexp(T x, int k);
template <> __device__ __forceinline__ float tldexp<float>(float x, int k) { return ldexpf(x, k); }
template <> __device__ __forceinline__ double tldexp<double>(double x, int k) { return ldexp(x, k); }
template <class T> __device__ __forceinline__ int texp(T x);  // x = m 2^e, 0.5 <= |m| < 1
template <> __device__ __forceinline__ int texp<float>(float x) { int e; frexpf(x, &e); return e; }
template <> __device__ __forceinline__ int texp<double>(double x) { int e; frexp(x, &e); return e; }

// Fixed-point accumulator of the P2G tile.
template <class T> struct AccOf;
template <> struct AccOf<float> {
  using U = uint32_t;
  static constexpr int BITS = 30;
  // LDS planes: two int32 sums per 64-bit word, (m v_x, m) and (m v_y, m v_z)
  static constexpr int PLANES = 2;
  // floor(x + 1/2) as int32 (two's complement), |x| < 2^31
  __device__ static U fix(float x) {
    int32_t r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return (U)r;
  }
  __device__ static float unfix(U q, int k) { return ldexpf((float)(int32_t)q, -k); }
  // hi 2^32 + lo as a 64-bit integer (lo sign-extended).  A 64-bit sum of
  // such words is (sum hi) 2^32 + (sum lo) exactly, so as long as both sums
  // fit in int32 (the per-item scales keep them below 2^31) unpack recovers
  // them: one 64-bit LDS atomic does the work of two 32-bit ones.
  __device__ static unsigned long long pack(float hi, float lo) {
    return ((unsigned long long)fix(hi) << 32) + (unsigned long long)(long long)(int32_t)fix(lo);
  }
  // the same word when lo >= 0 (a mass: non-negative weight times a positive
  // particle mass), without the sign carry: two VALU instructions fewer per node
  __device__ static unsigned long long pack_nonneg_lo(float hi, float lo) {
    return ((unsigned long long)fix(hi) << 32) | (unsigned long long)fix(lo);
  }
  __device__ static void unpack(unsigned long long w, int32_t &hi, int32_t &lo) {
    lo = (int32_t)(uint32_t)w;
    hi = (int32_t)(uint32_t)((w - (unsigned long long)(long long)lo) >> 32);
  }
};
template <> struct AccOf<double> {
  using U = unsigned long long;
  static constexpr int BITS = 62;
  static constexpr int PLANES = 4;  // one int64 sum per component
  // round(x) as int64 from exact steps, |x| < 2^62: |f| = hi 2^32 + lo with
  // hi < 2^30 and lo = |f| - hi 2^32 an integer below 2^32
  __device__ static U fix(double x) {
    const double f = rint(x);
    const double a = f < 0.0 ? -f : f;
    const double hi = floor(ldexp(a, -32));
    const double lo = a - hi * 4294967296.0;
    const U u = ((U)(uint32_t)hi << 32) + (U)(uint32_t)lo;
    return f < 0.0 ? (U)0 - u : u;
  }
  __device__ static double unfix(U q, int k) {
    const double d = (double)(int32_t)((long long)q >> 32) * 4294967296.0 + (double)(uint32_t)q;
    return ldexp(d, -k);
  }
};

template <class T> __device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Quantized records: at least 4 waves per SIMD (<= 128 VGPRs); measured
// ~11% faster than the 3 the register allocator picks by itself.
template <int ST> struct TiledWaves { static constexpr int W = ST == AQ_STORE_QUANT ? 4 : 1; };

template <int DIM, class T>
__device__ __forceinline__ void tile_origin(const int4 &item, int nb, int (&o)[3]) {
  int bc[3];
  block_coords<DIM>(item.x, nb, bc);
#pragma unroll
  for (int a = 0; a < 3; ++a) o[a] = a < DIM ? bc[a] * Tile<DIM>::E - 1 : 0;
}

// Loads and stages an item's grid-velocity tile node by node (a rolled
// loop keeps few VGPRs live).
template <int DIM, class T>
__device__ __forceinline__ void tile_copy(const ParamsT<T> &P, const int (&o)[3], int tid,
                                          const vec4_t<T> *__restrict__ grid_v, vec4_t<T> *lds) {
  using TL = Tile<DIM>;
  for (int nd = tid; nd < TL::TN; nd += kItem) {
    int t[3];
    TL::node_coords(nd, t);
    const int gi = o[0] + t[0], gj = o[1] + t[1], gk = o[2] + t[2];
    vec4_t<T> val;
    val.x = val.y = val.z = val.w = (T)0;
    const bool in = gi >= 0 && gi < P.n && gj >= 0 && gj < P.n && (DIM == 2 || (gk >= 0 && gk < P.n));
    if (in) val = grid_v[node_index<DIM>(gi, gj, gk, P.nb)];
    lds[TL::slot(t[0], t[1], t[2])] = val;
  }
}

// Everything the tiled kernels need of one simulation.
template <class T> struct TiledArgs {
  StateK s;
  CodecK cod;
  ParamsT<T> P;
  int64_t n;
  const int4 *items;
  const vec4_t<T> *grid_v;
  vec4_t<T> *tiles;
  vec4_t<T> *grid_f;
  uint32_t *fflags;
  int32_t *flist;  // blocks newly flagged by untiled scatters (null: flags only)
  uint32_t *fcnt;
  // per grid block, bit i: the block's item i (bit 31: items 31 and later,
  // which always write it) wrote its tile's drift margin; null: every item
  // writes its whole tile
  uint32_t *bmargin;
  const int32_t *item_off;  // first item of each block (with bmargin)
};

// Work of one simulation: its `count` items.
template <class T> struct OneSim {
  TiledArgs<T> a;
  int count;
  __device__ __forceinline__ int total() const { return count; }
  __device__ __forceinline__ const TiledArgs<T> &args(int it, int &li) const {
    li = it;
    return a;
  }
  __device__ __forceinline__ uint32_t key_base(uint32_t kb) const { return kb; }
};

// Work of a group of simulations: the items of all of them, numbered
// simulation after simulation (prefix[s] = items before simulation s,
// prefix[nsims] = all).  An item's simulation is found by binary search
// over the prefix; item numbers are uniform over a workgroup, so the search
// and the descriptor loads are scalar.  The descriptors' codec key is the
// seed's key; the dithering key of substep `step` is hash32(step ^ it), as
// for a single simulation.
template <class T> struct GroupSims {
  const TiledArgs<T> *__restrict__ descs;
  const int32_t *__restrict__ prefix;
  int32_t nsims;
  uint32_t step;
  __device__ __forceinline__ uint32_t key_base(uint32_t seed_key) const { return hash32(step ^ seed_key); }
  __device__ __forceinline__ int total() const { return prefix[nsims]; }
  __device__ __forceinline__ const TiledArgs<T> &args(int it, int &li) const {
    int lo = 0, hi = nsims;  // prefix[lo] <= it < prefix[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= it)
        lo = mid;
      else
        hi = mid;
    }
    li = it - prefix[lo];
    return descs[lo];
  }
};

// A workgroup walks its items (of one simulation, or of a group of them,
// `Src`); per item three barriers: the tile staged, the bounds of the
// fixed-point scales, the scatter complete.  (Loading the next item during
// the current item's scatter saves a barrier but holds ~20 more VGPRs: 2 %
// faster for the specialised quantized kernel with the particles of a block
// in random order, 1 % slower with the layered order; the fp32-storage
// kernel lost a wave per SIMD.)
template <int DIM, int MAT, int ST, int NW32, class T, bool DO_G2P, bool DO_P2G, class Src>
__device__ __forceinline__ void tiled_run(const Src &src) {
  using TL = Tile<DIM>;
  using ACC = AccOf<T>;
  using U = typename ACC::U;
  constexpr int TS = TL::TS, TN = TL::TN, TE = TL::TE, NC = Comps<DIM, MAT>::NC;
  __shared__ vec4_t<T> tv[DO_G2P ? TS : 1];
  constexpr int NP = ACC::PLANES;
  __shared__ unsigned long long ts[DO_P2G ? NP : 1][DO_P2G ? TS : 1];
  constexpr int NWAVES = kItem / 64;
  __shared__ T bpart[NWAVES][4];
  __shared__ uint32_t bdrift[NWAVES];
  __shared__ uint32_t rows[kItem * RowsOf<ST, NW32>::RS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  constexpr bool G2P_ONLY_LOAD = DO_G2P;  // G2P needs only the words of x and J/F
  uint32_t *row = rows + tid * RowsOf<ST, NW32>::RS;
  const int ni = src.total();
  // each XCD (blocks dealt round-robin to 8 XCDs) walks a contiguous item range
  const int xcd = (int)(blockIdx.x & 7u);
  const int per = (int)(gridDim.x >> 3);
  const int lo = (int)((int64_t)ni * xcd / 8), hi = (int)((int64_t)ni * (xcd + 1) / 8);
  int it = lo + (int)(blockIdx.x >> 3);
  if (it >= hi) return;  // uniform over the workgroup
  // LDS slot of each buffer position of the write-out, copied to LDS once
  // per workgroup (visible after the first barrier): the write-out then
  // looks it up with an LDS read instead of a global load per item
  __shared__ uint16_t wsl[DO_P2G ? TN : 1];
  if constexpr (DO_P2G)
    for (int q = tid; q < TN; q += kItem) wsl[q] = (uint16_t)tile_slot_at<DIM>(q);
  if constexpr (DO_P2G) {
    for (int q = tid; q < TS; q += kItem)
#pragma unroll
      for (int a = 0; a < NP; ++a) ts[a][q] = 0ull;  // cleared before the first barrier
  }
  for (; it < hi; it += per) {
    int li;  // the item's number in its simulation
    const TiledArgs<T> &A = src.args(it, li);
    const StateK &s = A.s;
    CodecK cod = A.cod;
    cod.key_base = src.key_base(cod.key_base);
    const ParamsT<T> &P = A.P;
    const int64_t n = A.n;
    const vec4_t<T> *__restrict__ grid_v = A.grid_v;
    vec4_t<T> *__restrict__ grid_f = A.grid_f;
    uint32_t *__restrict__ fflags = A.fflags;
#ifdef AQ_JIT_LAYOUT
    constexpr uint32_t g2p_mask = jit::kG2PMask;
#else
    const uint32_t g2p_mask = s.g2p_mask;
#endif
    const uint32_t load_mask = DO_G2P ? g2p_mask : 0xFFFFFFFFu;
    const int4 item = A.items[li];
    int o[3];
    tile_origin<DIM, T>(item, P.nb, o);
    // the particle's loads go out before the tile's, so their latencies overlap
    Fetched<ST, NW32> fw;
    T v[NC];
    if (tid < item.z && (int64_t)item.y + tid < n)
      fetch_state<DIM, ST, NW32, NC, T, G2P_ONLY_LOAD>(s, (int64_t)item.y + tid, load_mask, fw, v);
    if constexpr (DO_G2P) tile_copy<DIM, T>(P, o, tid, grid_v, tv);
    __syncthreads();  // tile staged; the previous item's tile written out
    const bool has = tid < item.z && (int64_t)item.y + tid < n;
    const int64_t p = (int64_t)item.y + tid;
    T out[NC];
    if (has) {
      if constexpr (DO_G2P) {
        finish_state<DIM, ST, NW32, NC, T, true>(s, cod, g2p_mask, fw, row, v);
        const Stencil<DIM, T> S = stencil_of<DIM, MAT, T>(v, P);
        int rel[3] = {0, 0, 0};
        bool inside = true;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          rel[a] = S.base[a] - o[a];
          inside = inside && rel[a] >= 0 && rel[a] + 2 < TE;
        }
        if (inside) {
          g2p_apply<DIM, MAT, T>(
              v, P, S, [&](int i, int j, int k) { return tv[TL::slot(rel[0] + i, rel[1] + j, rel[2] + k)]; }, out);
        } else {
          g2p_particle<DIM, MAT, T>(v, P, grid_v, out);
        }
        store_state<ST, NW32, NC, T>(s, cod, p, out, row);  // out: what storage now holds
      } else {
        finish_state<DIM, ST, NW32, NC, T>(s, cod, 0xFFFFFFFFu, fw, row, v);
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c] = v[c];
      }
    }
    Stencil<DIM, T> S;
    P2GTerms<DIM, T> R;
    int rel[3] = {0, 0, 0};
    bool inside = false;
    if constexpr (DO_P2G) {
      T B[4] = {(T)0, (T)0, (T)0, (T)0};
      bool drift = false;  // the stencil reaches the tile's drift margin
      if (has) {
        S = stencil_of<DIM, MAT, T>(out, P);
        R = p2g_terms<DIM, MAT, T>(out, P, S);
        inside = true;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          rel[a] = S.base[a] - o[a];
          // fx outside [0.5, 1.5] (a base clamped at the domain edge) can make
          // a weight negative, and the packed mass sums need them >= 0: such a
          // particle scatters through the float-atomic path
          inside = inside && rel[a] >= 0 && rel[a] + 2 < TE && S.fx[a] >= (T)0.5 && S.fx[a] <= (T)1.5;
          drift = drift || rel[a] == 0 || rel[a] == TE - 3;
        }
        if (inside) p2g_bounds<DIM, T>(R, S, B);
        drift = drift && inside;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) B[c] = wave_sum<T>(B[c]);
      const bool wdrift = __ballot(drift) != 0ull;
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) bpart[wave][c] = B[c];
        bdrift[wave] = wdrift ? 1u : 0u;
      }
    }
    __syncthreads();  // item bounds complete; G2P done with the tile
    int k[4] = {0, 0, 0, 0};
    bool margin = true;  // the write-out covers the drift margin
    if constexpr (DO_P2G) {
      if (A.bmargin) {
        uint32_t d = 0u;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) d |= bdrift[w];
        const int ib = li - A.item_off[item.x];  // the item's number in its block
        margin = d != 0u || ib >= 31;
        if (margin && tid == 0) atomicOr(A.bmargin + item.x, 1u << (ib < 31 ? ib : 31));
      }
      bool finite = true;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T sum = bpart[0][c];
#pragma unroll
        for (int w = 1; w < NWAVES; ++w) sum += bpart[w][c];
        finite = finite && sum <= (T)1.0e37;
        k[c] = ACC::BITS - texp<T>(sum * ((T)1 + (T)1 / (T)1024));
      }
      if (has) {
        if (inside && finite) {
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            R.c[a] = tldexp<T>(R.c[a], k[a]);
#pragma unroll
            for (int b = 0; b < DIM; ++b) R.A[a][b] = tldexp<T>(R.A[a][b], k[a]);
          }
          R.pm = tldexp<T>(R.pm, k[3]);
          p2g_emit<DIM, T>(R, S, [&](int i, int j, int kk, const T (&mom)[3], T m) {
            const int q = TL::slot(rel[0] + i, rel[1] + j, rel[2] + kk);
            if constexpr (NP == 2) {
              atomicAdd(&ts[0][q], ACC::pack_nonneg_lo(mom[0], m));
              atomicAdd(&ts[1][q], ACC::pack(mom[1], DIM == 3 ? mom[2] : (T)0));
            } else {
#pragma unroll
              for (int a = 0; a < DIM; ++a) atomicAdd(&ts[a][q], (unsigned long long)ACC::fix(mom[a]));
              atomicAdd(&ts[3][q], (unsigned long long)ACC::fix(m));
            }
          });
        } else {
          p2g_particle<DIM, MAT, T>(out, P, grid_f, FlagSink{fflags, A.flist, A.fcnt});
        }
      }
    }
    if constexpr (DO_P2G) {
      __syncthreads();  // tile sums complete
      vec4_t<T> *dst = A.tiles + (int64_t)li * TN;
      for (int nd = tid; nd < TN; nd += kItem) {  // nd: position in the buffer
        const int ws = wsl[nd];
        // without drift the margin's LDS sums are zero (never added to, so
        // they need no clearing either) and the gather does not read them
        if (!margin && (ws & kMarginBit)) continue;
        const int q = ws & (kMarginBit - 1);
        vec4_t<T> val;
        if constexpr (NP == 2) {
          int32_t mx, m, my, mz;
          ACC::unpack(ts[0][q], mx, m);
          ACC::unpack(ts[1][q], my, mz);
          val.x = ACC::unfix((U)mx, k[0]);
          val.y = ACC::unfix((U)my, k[1]);
          val.z = DIM == 3 ? ACC::unfix((U)mz, k[2]) : (T)0;
          val.w = ACC::unfix((U)m, k[3]);
        } else {
          val.x = ACC::unfix(ts[0][q], k[0]);
          val.y = ACC::unfix(ts[1][q], k[1]);
          val.z = DIM == 3 ? ACC::unfix(ts[2][q], k[2]) : (T)0;
          val.w = ACC::unfix(ts[3][q], k[3]);
        }
        dst[nd] = val;
#pragma unroll
        for (int a = 0; a < NP; ++a) ts[a][q] = 0ull;
      }
    }
  }
}

// One simulation's items (the count is the last sort's, read on the device).
template <int DIM, int MAT, int ST, int NW32, class T, bool DO_G2P, bool DO_P2G>
__device__ __forceinline__ void tiled_body(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                           const int4 *__restrict__ items, const int32_t *__restrict__ nitems,
                                           const vec4_t<T> *__restrict__ grid_v, vec4_t<T> *__restrict__ tiles,
                                           vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags,
                                           int32_t *flist, uint32_t *gcnt, uint32_t ctl, uint32_t *bmargin) {
  // commit of the last list gather's appends (no thread of this kernel
  // reads gcnt[0] or the cleared counter; appends go to the other one)
  const uint32_t commit = (ctl >> 2) & 3u;
  if (commit && blockIdx.x == 0 && threadIdx.x == 0) {
    gcnt[0] = gcnt[3];
    gcnt[commit] = 0u;
  }
  uint32_t *fcnt = flist ? gcnt + (ctl & 3u) : nullptr;
  const int32_t nblocks = DIM == 3 ? P.nb * P.nb * P.nb : P.nb * P.nb;
  const OneSim<T> src{{s, cod, P, n, items, grid_v, tiles, grid_f, fflags, flist, fcnt, bmargin, nitems - nblocks},
                      *nitems};
  tiled_run<DIM, MAT, ST, NW32, T, DO_G2P, DO_P2G>(src);
}

template <int DIM, int MAT, int ST, int NW32, class T, bool DO_G2P, bool DO_P2G>
__global__ void __launch_bounds__(kItem) __attribute__((amdgpu_waves_per_eu(TiledWaves<ST>::W)))
k_tiled(StateK s, CodecK cod, ParamsT<T> P, int64_t n, const int4 *__restrict__ items,
        const int32_t *__restrict__ nitems, const vec4_t<T> *__restrict__ grid_v, vec4_t<T> *__restrict__ tiles,
        vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags, int32_t *flist, uint32_t *gcnt,
        uint32_t ctl, uint32_t *bmargin) {
  tiled_body<DIM, MAT, ST, NW32, T, DO_G2P, DO_P2G>(s, cod, P, n, items, nitems, grid_v, tiles, grid_f, fflags,
                                                    flist, gcnt, ctl, bmargin);
}

// The fused G2P -> P2G substep of a group of quantized simulations (one
// dimension and material, any record layouts) in one launch.
template <int DIM, int MAT, int NW32>
__global__ void __launch_bounds__(kItem) __attribute__((amdgpu_waves_per_eu(TiledWaves<AQ_STORE_QUANT>::W)))
k_tiled_group(const TiledArgs<float> *__restrict__ descs, const int32_t *__restrict__ prefix, int32_t nsims,
              uint32_t step) {
  const GroupSims<float> src{descs, prefix, nsims, step};
  tiled_run<DIM, MAT, AQ_STORE_QUANT, NW32, float, true, true>(src);
}

}  // namespace aq
