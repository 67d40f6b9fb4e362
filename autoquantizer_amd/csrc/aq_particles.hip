// libaqmpm: the separate P2G and G2P launches (tiled and untiled) of every
// dimension, material and storage; a translation unit of its own so that the
// library's kernels compile in parallel.
#include "aq_sim.hpp"
#include "aq_untiled.hpp"

// P2G of `state` (tiled when the last sort describes it); *tiled tells which.
int launch_p2g(aq_sim *s, const void *state, int64_t n, bool *tiled, hipStream_t st) {
  *tiled = n > 0 && items_valid(s, state, n);
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state, nullptr);
  CodecK cod = make_codec(s, 0u);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    if (*tiled)
      hipLaunchKernelGGL((k_tiled<DIM, MAT, ST, NW32, T, false, true>), dim3(tblocks(s, n)), dim3(kItem), 0, st, sk,
                         cod, params_of<T>(s), n, s->items, s->item_off + s->nblocks, (const vec4_t<T> *)nullptr,
                         (vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags, s->flist, s->gcnt + 1);
    else
      hipLaunchKernelGGL((k_p2g<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                         params_of<T>(s), n, (vec4_t<T> *)s->grid_f, s->fflags);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

int launch_g2p(aq_sim *s, void *state, int64_t n, const int32_t *ids, uint32_t step, hipStream_t st) {
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state, ids);
  CodecK cod = make_codec(s, step);
  const bool tiled = items_valid(s, state, n);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    if (tiled)
      hipLaunchKernelGGL((k_tiled<DIM, MAT, ST, NW32, T, true, false>), dim3(tblocks(s, n)), dim3(kItem), 0, st, sk,
                         cod, params_of<T>(s), n, s->items, s->item_off + s->nblocks,
                         (const vec4_t<T> *)s->grid_v, (vec4_t<T> *)nullptr, (vec4_t<T> *)nullptr,
                         (uint32_t *)nullptr, (int32_t *)nullptr, (uint32_t *)nullptr);
    else
      hipLaunchKernelGGL((k_g2p<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                         params_of<T>(s), n, (const vec4_t<T> *)s->grid_v);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}
