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
// in Tile; in the tile buffer, the core nodes grouped by the grid block they
// belong to, then the rim nodes, make_tile_order).  The rim (the drift
// margin) is written only when some particle of the item reached it: the
// item's rim flags (one bit per face) follow its nodes in the buffer.
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
  // Per axis the tile covers the last node of the previous block (t = 0),
  // the E nodes of its own block (1..E), the first 3 of the next (E+1..E+3).
  // A particle sorted into the block has its stencil base in the block, so
  // its stencil covers t = 1..E+2: the core.  Only a particle that drifted a
  // cell since the sort reaches t = 0 or t = E+3: the rim.
  static constexpr int NCORE = DIM == 3 ? (E + 2) * (E + 2) * (E + 2) : (E + 2) * (E + 2);
  // Nodes per item slot of the tile buffer: the TN tile nodes, then the
  // item's rim flags (in node TN), padded to 128 bytes.
  static constexpr int STRIDE = TN + 8;
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
  // Rim faces of tile node t: bit 2a when t_a = 0, bit 2a+1 when t_a = E+3.
  // The node can be non-zero only if the item's particles reached every
  // one of these faces.
  __host__ __device__ static constexpr uint32_t need(int x, int y, int z) {
    const int t[3] = {x, y, z};
    uint32_t m = 0u;
    for (int a = 0; a < DIM; ++a) {
      if (t[a] == 0) m |= 1u << (2 * a);
      if (t[a] == E + 3) m |= 2u << (2 * a);
    }
    return m;
  }
  // Position of core node (x, y, z) in its item's slot of the tile buffer.
  // Core nodes come first (positions 0..NCORE-1): per axis the core splits
  // into the E nodes of the own block and the 2 of the next, and the core
  // is stored as these 2^DIM groups (in group order), each row-major, which
  // is the gathering wavefront's lane order: a wavefront reads each covering
  // tile's core as one contiguous run.  Rim nodes follow (make_tile_order).
  __host__ __device__ static constexpr int core_pos(int x, int y, int z) {
    const int t[3] = {x, y, z};
    int g[3] = {0, 0, 0}, u[3] = {0, 0, 0}, sz[3] = {1, 1, 1};
    for (int a = 0; a < DIM; ++a) {
      g[a] = t[a] <= E ? 0 : 1;
      u[a] = t[a] <= E ? t[a] - 1 : t[a] - E - 1;
      sz[a] = g[a] == 0 ? E : 2;
    }
    const int cum[2] = {0, E};  // core nodes per axis before group g
    constexpr int C = E + 2;
    if constexpr (DIM == 3) {
      const int before = cum[g[0]] * C * C + sz[0] * cum[g[1]] * C + sz[0] * sz[1] * cum[g[2]];
      return before + (u[0] * sz[1] + u[1]) * sz[2] + u[2];
    } else {
      const int before = cum[g[0]] * C + sz[0] * cum[g[1]];
      return before + u[0] * sz[1] + u[1];
    }
  }
};

// Per position of an item's slot in the tile buffer (core nodes at
// Tile::core_pos, then the rim nodes in row-major order): the LDS slot of its node in bits 0..9 and its rim faces
// (Tile::need) in bits 10..15; and per tile node (x TE + y) TE + z its
// position.  Built at compile time: the write-out walks the slot in order,
// so each wavefront stores a contiguous run, and the gather looks the
// positions up.
template <int DIM> struct TileOrder {
  uint16_t at[Tile<DIM>::TN];
  uint16_t pos[Tile<DIM>::TN];
};
template <int DIM> constexpr TileOrder<DIM> make_tile_order() {
  using TL = Tile<DIM>;
  static_assert(TL::TS <= 1024, "LDS slots fit 10 bits");
  TileOrder<DIM> o{};
  int rim = TL::NCORE;  // rim nodes in row-major order after the core
  for (int x = 0; x < TL::TE; ++x)
    for (int y = 0; y < TL::TE; ++y)
      for (int z = 0; z < (DIM == 3 ? TL::TE : 1); ++z) {
        const uint32_t need = TL::need(x, y, z);
        const int q = need ? rim++ : TL::core_pos(x, y, z);
        o.at[q] = (uint16_t)(TL::slot(x, y, z) | (need << 10));
        o.pos[DIM == 3 ? (x * TL::TE + y) * TL::TE + z : x * TL::TE + y] = (uint16_t)q;
      }
  return o;
}
__constant__ constexpr TileOrder<3> kTileOrder3 = make_tile_order<3>();
__constant__ constexpr TileOrder<2> kTileOrder2 = make_tile_order<2>();
// LDS slot and rim faces of the node at position `pos` of a slot.
template <int DIM> __device__ __forceinline__ int tile_slot_at(int pos) {
  if constexpr (DIM == 3)
    return kTileOrder3.at[pos];
  else
    return kTileOrder2.at[pos];
}
// Position of tile node (x, y, z) in its slot.
template <int DIM> __device__ __forceinline__ int tile_pos(int x, int y, int z) {
  if constexpr (DIM == 3)
    return kTileOrder3.pos[(x * Tile<3>::TE + y) * Tile<3>::TE + z];
  else
    return kTileOrder2.pos[x * Tile<2>::TE + y];
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

// A workgroup walks its items; per item three barriers: the tile staged,
// the bounds of the fixed-point scales, the scatter complete.  (Loading the
// next item during the current item's scatter saves a barrier but holds
// ~20 more VGPRs: 2 % faster for the specialised quantized kernel with the
// particles of a block in random order, 1 % slower with the layered order;
// the fp32-storage kernel lost a wave per SIMD.)
template <int DIM, int MAT, int ST, int NW32, class T, bool DO_G2P, bool DO_P2G>
__device__ __forceinline__ void tiled_body(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                           const int4 *__restrict__ items, const int32_t *__restrict__ nitems,
                                           const vec4_t<T> *__restrict__ grid_v, vec4_t<T> *__restrict__ tiles,
                                           vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags) {
  using TL = Tile<DIM>;
  using ACC = AccOf<T>;
  using U = typename ACC::U;
  constexpr int TS = TL::TS, TN = TL::TN, TE = TL::TE, NC = Comps<DIM, MAT>::NC;
  __shared__ vec4_t<T> tv[DO_G2P ? TS : 1];
  constexpr int NP = ACC::PLANES;
  __shared__ unsigned long long ts[DO_P2G ? NP : 1][DO_P2G ? TS : 1];
  constexpr int NWAVES = kItem / 64;
  __shared__ T bpart[NWAVES][4];
  // rim faces reached by the item's particles, two buffers by item parity
  // (one is cleared while the other collects)
  __shared__ uint32_t rimfl[2];
  __shared__ uint32_t rows[kItem * RowsOf<ST, NW32>::RS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#ifdef AQ_JIT_LAYOUT
  constexpr uint32_t g2p_mask = jit::kG2PMask;
#else
  const uint32_t g2p_mask = s.g2p_mask;
#endif
  constexpr bool G2P_ONLY_LOAD = DO_G2P;  // G2P needs only the words of x and J/F
  const uint32_t load_mask = DO_G2P ? g2p_mask : 0xFFFFFFFFu;
  uint32_t *row = rows + tid * RowsOf<ST, NW32>::RS;
  const int ni = *nitems;
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
    if (tid < 2) rimfl[tid] = 0u;
  }
  for (uint32_t par = 0u; it < hi; it += per, par ^= 1u) {
    const int4 item = items[it];
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
        }
        if (inside) p2g_bounds<DIM, T>(R, S, B);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) B[c] = wave_sum<T>(B[c]);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) bpart[wave][c] = B[c];
      }
    }
    __syncthreads();  // item bounds complete; G2P done with the tile
    int k[4] = {0, 0, 0, 0};
    if constexpr (DO_P2G) {
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
          // faces of the tile's rim this stencil reaches (it covers rel..rel+2)
          uint32_t fl = 0u;
#pragma unroll
          for (int a = 0; a < DIM; ++a)
            fl |= (rel[a] == 0 ? 1u : 0u) << (2 * a) | (rel[a] == TE - 3 ? 2u : 0u) << (2 * a);
          if (fl) atomicOr(&rimfl[par], fl);
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
          p2g_particle<DIM, MAT, T>(out, P, grid_f, fflags);
        }
      }
    }
    if constexpr (DO_P2G) {
      __syncthreads();  // tile sums complete
      const uint32_t fl = rimfl[par];
      if (tid == 0) rimfl[par ^ 1u] = 0u;  // read by every thread before the previous barrier
      vec4_t<T> *dst = tiles + (int64_t)it * TL::STRIDE;
      if (tid == 0) *(uint32_t *)(dst + TN) = fl;
      // the core always; rim nodes only when the item reached their faces
      // (untouched rim nodes are zero in LDS and need no clearing)
      const int lim = fl ? TN : TL::NCORE;
      for (int nd = tid; nd < lim; nd += kItem) {  // nd: position in the buffer
        const uint32_t sn = wsl[nd];
        const int q = (int)(sn & 1023u);
        const uint32_t need = sn >> 10;
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
        if ((fl & need) == need) dst[nd] = val;
#pragma unroll
        for (int a = 0; a < NP; ++a) ts[a][q] = 0ull;
      }
    }
  }
}

template <int DIM, int MAT, int ST, int NW32, class T, bool DO_G2P, bool DO_P2G>
__global__ void __launch_bounds__(kItem) __attribute__((amdgpu_waves_per_eu(TiledWaves<ST>::W)))
k_tiled(StateK s, CodecK cod, ParamsT<T> P, int64_t n, const int4 *__restrict__ items,
        const int32_t *__restrict__ nitems, const vec4_t<T> *__restrict__ grid_v, vec4_t<T> *__restrict__ tiles,
        vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags) {
  tiled_body<DIM, MAT, ST, NW32, T, DO_G2P, DO_P2G>(s, cod, P, n, items, nitems, grid_v, tiles, grid_f, fflags);
}

}  // namespace aq
