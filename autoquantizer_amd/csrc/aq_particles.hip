// libaqmpm: the separate P2G and G2P launches (tiled and untiled) of every
// dimension, material and storage; a translation unit of its own so that the
// library's kernels compile in parallel.
#include "aq_sim.hpp"
#include "aq_untiled.hpp"

// P2G of `state` (tiled when the last sort describes it); *tiled tells which.
// whole: every tile written whole (for a gather that reads them so), else
// drift margins only where the next list gather's flags say.
int launch_p2g(aq_sim *s, const void *state, int64_t n, bool *tiled, hipStream_t st, bool whole) {
  *tiled = n > 0 && items_valid(s, state, n);
  if (n == 0) return AQ_OK;
  StateK sk = make_state(s, state, nullptr);
  CodecK cod = make_codec(s, 0u);
  uint32_t *bm = whole ? nullptr : s->bmargin + (int64_t)s->list_parity * s->nblocks;  // the next gather's flags
  uint32_t ctl = *tiled ? take_list_ctl(s) : 0u;
  if (*tiled && s->jit_fn_p2g) {
    ParamsT<float> P = s->pf;
    const int4 *items = s->items;
    const int32_t *nitems = s->item_off + s->nblocks;
    const void *gv = nullptr;
    void *tl = s->tiles, *gf = s->grid_f;
    uint32_t *ff = s->fflags;
    int32_t *fl = s->flist;
    uint32_t *gc = s->gcnt;
    void *args[] = {&sk, &cod, &P, &n, &items, &nitems, &gv, &tl, &gf, &ff, &fl, &gc, &ctl, &bm};
    HIP_TRY(hipModuleLaunchKernel(s->jit_fn_p2g, tblocks(s, n), 1, 1, kItem, 1, 1, 0, st, args, nullptr));
    return AQ_OK;
  }
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    if (*tiled)
      hipLaunchKernelGGL((k_tiled<DIM, MAT, ST, NW32, T, false, true>), dim3(tblocks(s, n)), dim3(kItem), 0, st, sk,
                         cod, params_of<T>(s), n, s->items, s->item_off + s->nblocks, (const vec4_t<T> *)nullptr,
                         (vec4_t<T> *)s->tiles, (vec4_t<T> *)s->grid_f, s->fflags, s->flist, s->gcnt, ctl, bm);
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
  uint32_t ctl = tiled ? take_list_ctl(s) : 0u;  // G2P only: carries a commit, appends nothing
  if (tiled && s->jit_fn_g2p) {
    ParamsT<float> P = s->pf;
    const int4 *items = s->items;
    const int32_t *nitems = s->item_off + s->nblocks;
    const void *gv = s->grid_v;
    void *tl = nullptr, *gf = nullptr;
    uint32_t *ff = nullptr, *bm = nullptr;
    int32_t *fl = nullptr;
    uint32_t *gc = s->gcnt;
    void *args[] = {&sk, &cod, &P, &n, &items, &nitems, &gv, &tl, &gf, &ff, &fl, &gc, &ctl, &bm};
    HIP_TRY(hipModuleLaunchKernel(s->jit_fn_g2p, tblocks(s, n), 1, 1, kItem, 1, 1, 0, st, args, nullptr));
    return AQ_OK;
  }
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    if (tiled)
      hipLaunchKernelGGL((k_tiled<DIM, MAT, ST, NW32, T, true, false>), dim3(tblocks(s, n)), dim3(kItem), 0, st, sk,
                         cod, params_of<T>(s), n, s->items, s->item_off + s->nblocks,
                         (const vec4_t<T> *)s->grid_v, (vec4_t<T> *)nullptr, (vec4_t<T> *)nullptr,
                         (uint32_t *)nullptr, (int32_t *)nullptr, s->gcnt, ctl, (uint32_t *)nullptr);
    else
      hipLaunchKernelGGL((k_g2p<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                         params_of<T>(s), n, (const vec4_t<T> *)s->grid_v);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

// ---- records <-> values in the simulation's compute type -------------------
template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_sim_pack(StateK s, CodecK cod, const T *__restrict__ vals, int64_t n) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = vals[(int64_t)c * n + p];
  store_state<ST, NW32, NC, T>(s, cod, p, v, rows + threadIdx.x * RowsOf<ST, NW32>::RS);
}

template <int DIM, int MAT, int ST, int NW32, class T>
__global__ void __launch_bounds__(256) k_sim_unpack(StateK s, CodecK cod, T *__restrict__ vals, int64_t n) {
  __shared__ uint32_t rows[256 * RowsOf<ST, NW32>::RS];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  constexpr int NC = Comps<DIM, MAT>::NC;
  T v[NC];
  load_state<DIM, ST, NW32, NC, T>(s, cod, p, 0xFFFFFFFFu, rows + threadIdx.x * RowsOf<ST, NW32>::RS, v);
#pragma unroll
  for (int c = 0; c < NC; ++c) vals[(int64_t)c * n + p] = v[c];
}

extern "C" int aq_sim_pack(aq_sim *s, const void *vals, int64_t n, uint32_t step, uint64_t *words, void *stream) {
  int rc = check_n(s, words, n);
  if (rc) return rc;
  if (s->d.storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "records belong to quantized storage");
  if (n && !vals) return fail(AQ_ERR_INVALID, "null values");
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  const StateK sk = make_state(s, words, nullptr);
  const CodecK cod = make_codec(s, step);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_sim_pack<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                       (const T *)vals, n);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_sim_unpack(aq_sim *s, const uint64_t *words, int64_t n, void *vals, void *stream) {
  int rc = check_n(s, words, n);
  if (rc) return rc;
  if (s->d.storage != AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "records belong to quantized storage");
  if (n && !vals) return fail(AQ_ERR_INVALID, "null values");
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  const StateK sk = make_state(s, words, nullptr);
  const CodecK cod = make_codec(s, 0u);
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    hipLaunchKernelGGL((k_sim_unpack<DIM, MAT, ST, NW32, T>), dim3(pblocks(n)), dim3(256), 0, st, sk, cod,
                       (T *)vals, n);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}
