// libaqmpm: the adjoint of the plain substep (aq_adj_*).
#include "aq_sim.hpp"
#include "aq_adjoint_kernels.hpp"

// ---------------------------------------------------------------------------
// adjoint of the plain substep
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_zero_active(int32_t nblocks, const uint32_t *__restrict__ act,
                                                     vec4_t<T> *__restrict__ g) {
  const int32_t b = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
  if (b >= nblocks || !act[b]) return;
  vec4_t<T> z;
  z.x = z.y = z.z = z.w = (T)0;
  g[(int64_t)b * 64 + (threadIdx.x & 63)] = z;
}

extern "C" int64_t aq_adj_workspace_bytes(const aq_sim *s) {
  if (!s) return -1;
  return 2 * s->nodes * s->node_bytes;
}

extern "C" int aq_adj_substep(aq_sim *s, uint32_t step, const void *state, int64_t n, const void *adj_next,
                              void *adj_out, double *gsum, void *ws, int64_t ws_bytes, void *stream) {
  int rc = check_n(s, state, n);
  if (rc) return rc;
  if (s->d.storage == AQ_STORE_QUANT) return fail(AQ_ERR_INVALID, "the adjoint runs on plain (f32/f64) storage");
  if (s->d.bound_clip) return fail(AQ_ERR_INVALID, "the adjoint of the bound clip is not implemented");
  if (n && (!adj_next || !adj_out)) return fail(AQ_ERR_INVALID, "null adjoint buffer");
  if (!ws || ws_bytes < aq_adj_workspace_bytes(s)) return fail(AQ_ERR_CAPACITY, "adjoint workspace too small");
  if (n == 0) return AQ_OK;
  hipStream_t st = (hipStream_t)stream;
  void *pre = ws;
  void *gbar = (char *)ws + s->nodes * s->node_bytes;
  // forward recompute of s_t's grid: P2G (tiled when the last sort describes
  // `state`), gather, copy of (m v, m), grid update
  bool tiled;
  if ((rc = launch_p2g(s, state, n, &tiled, st))) return rc;
  if ((rc = launch_gather(s, tiled, false, 0u, st))) return rc;
  if ((rc = settle_list(s, st))) return rc;
  HIP_TRY(hipMemcpyAsync(pre, s->grid_v, (size_t)(s->nodes * s->node_bytes), hipMemcpyDeviceToDevice, st));
  if ((rc = aq_grid_op(s, step, stream))) return rc;
  const ColliderK K = collider_at(s, step);
  const int64_t cap = s->d.capacity;
  dispatch(s, [&](auto tag_) {
    TAG_TYPES(tag_)
    (void)ST;
    (void)NW32;
    const ParamsT<T> &P = params_of<T>(s);
    hipLaunchKernelGGL((k_zero_active<T>), dim3(gblocks(s)), dim3(256), 0, st, s->nblocks, (const uint32_t *)s->act,
                       (vec4_t<T> *)gbar);
    hipLaunchKernelGGL((k_adj_g2p<DIM, MAT, T>), dim3(pblocks(n)), dim3(256), 0, st, (const T *)state,
                       (const T *)adj_next, (T *)adj_out, cap, n, P, (const vec4_t<T> *)s->grid_v, (T *)gbar);
    hipLaunchKernelGGL((k_adj_grid<DIM, T>), dim3(gblocks(s)), dim3(256), 0, st, P, K, s->nblocks,
                       (const uint32_t *)s->act, (const vec4_t<T> *)pre, (T *)gbar);
    hipLaunchKernelGGL((k_adj_p2g<DIM, MAT, T>), dim3(pblocks(n)), dim3(256), 0, st, (const T *)state, (T *)adj_out,
                       cap, n, P, (const T *)gbar, gsum);
  });
  LAUNCH_CHECK();
  return AQ_OK;
}
