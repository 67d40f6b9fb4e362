// libaqmpm: the fused G2P -> P2G launch (tiled, specialised or untiled) of
// every dimension, material and storage; a translation unit of its own so
// that the library's kernels compile in parallel.
#include "aq_sim.hpp"
#include "aq_untiled.hpp"

// G2P of substep `step` fused with the next P2G; *tiled tells which P2G ran.
int launch_g2p2g(aq_sim *s, void *state, int64_t n, const int32_t *ids, uint32_t step, bool *tiled,
                        hipStream_t st) {
  *tiled = n > 0 && items_valid(s, state, n);
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state, ids);
  CodecK cod = make_codec(s, step);
  if (*tiled && s->jit_fn) {
    ParamsT<float> P = s->pf;
    const int4 *items = s->items;
    const int32_t *nitems = s->item_off + s->nblocks;
    const void *gv = s->grid_v;
    void *tl = s->tiles, *gf = s->grid_f;
    uint32_t *ff = s->fflags;
    int32_t *fl = s->flist;
    uint32_t *gc = s->gcnt;
    uint32_t *bm = s->bmargin + (int64_t)s->list_parity * s->nblocks;
    uint32_t ctl = take_list_ctl(s);
    void *args[] = {&sk, &cod, &P, &n, &items, &nitems, &gv, &tl, &gf, &ff, &fl, &gc, &ctl, &bm};
    HIP_TRY(hipModuleLaunchKernel(s->jit_fn, tblocks(s, n), 1, 1, kItem, 1, 1, 0, st, args, nullptr));
    return AQ_OK;
  }
  uint32_t *bm = s->bmargin + (int64_t)s->list_parity * s->nblocks;  // the next gather's flags
  const uint32_t ctl = *tiled ? take_list_ctl(s) : 0u;
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    if (*tiled)
      hipLaunchKernelGGL((k_tiled<DIM, MAT, ST, NW32, T, true, true>), dim3(tblocks(s, n)), dim3(kItem), 0, st, sk,
                         cod, params_of<T>(s), n, s->items, s->item_off + s->nblocks,
                         (const vec4_t<T> *)s->grid_v, (vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags,
                         s->flist, s->gcnt, ctl, bm);
    else
      hipLaunchKernelGGL((k_g2p2g<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                         params_of<T>(s), n, (const vec4_t<T> *)s->grid_v, (vec4_t<T> *)s->grid_f, s->fflags);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}
