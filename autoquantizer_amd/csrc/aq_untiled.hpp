// Untiled MPM kernels of libaqmpm (included by aq_particles.hip and
// aq_fused.hip).
#pragma once
#include "aq_sim.hpp"

// Untiled kernels (one particle per thread, grid in global memory).  P2G
// scatters with float atomics into the accumulation grid grid_f and flags
// the touched blocks; the gather kernel turns grid_f into grid_v.
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_p2g(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                             vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC];
  load_state<DIM, ST, NW32, NC, T>(s, cod, p, 0xFFFFFFFFu, rows + threadIdx.x * RowsOf<ST, NW32>::RS, v);
  p2g_particle<DIM, MAT, T>(v, P, grid_f, FlagSink{fflags, nullptr, nullptr});
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_g2p(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                             const vec4_t<T> *__restrict__ grid_v) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  uint32_t *row = rows + threadIdx.x * RowsOf<ST, NW32>::RS;
  T v[NC], out[NC];
  load_state<DIM, ST, NW32, NC, T, true>(s, cod, p, s.g2p_mask, row, v);
  g2p_particle<DIM, MAT, T>(v, P, grid_v, out);
  store_state<ST, NW32, NC, T>(s, cod, p, out, row);
}

// G2P of substep t (reading grid_v) fused with P2G of substep t+1
// (scattering into grid_f): the state is read once and written once.
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_g2p2g(StateK s, CodecK cod, ParamsT<T> P, int64_t n,
                                               const vec4_t<T> *__restrict__ grid_v,
                                               vec4_t<T> *__restrict__ grid_f, uint32_t *__restrict__ fflags) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  uint32_t *row = rows + threadIdx.x * RowsOf<ST, NW32>::RS;
  T v[NC], out[NC];
  load_state<DIM, ST, NW32, NC, T, true>(s, cod, p, s.g2p_mask, row, v);
  g2p_particle<DIM, MAT, T>(v, P, grid_v, out);
  store_state<ST, NW32, NC, T>(s, cod, p, out, row);  // out: what storage now holds
  p2g_particle<DIM, MAT, T>(out, P, grid_f, FlagSink{fflags, nullptr, nullptr});
}
