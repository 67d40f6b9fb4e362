/*
 * aq_mpm.h -- C ABI of libaqmpm.so, the MI355X (gfx950) implementation of the
 * quantized MLS-MPM substep of Hanke98/AutoQuantizer.
 *
 * Paths in citations are relative to the reference root; TS/ is
 * taichi_submission/ (the reference's patched Taichi v0.8.4).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative AQ_ERR_* code; the
 *    message of the last error on the calling thread is aq_last_error().
 *    No C++ exception crosses the ABI.
 *  - Device pointers are caller-owned (the Python layer passes torch tensor
 *    storage); the library never allocates device memory.  A simulation's
 *    grid, flags and sort buffers live in a caller-provided workspace of
 *    aq_sim_workspace_bytes() bytes.
 *  - `stream` is a hipStream_t (NULL = default stream).  Nothing synchronises
 *    implicitly; work is enqueued on `stream` in call order.
 *  - Particle state is stored as word planes: plane w of a record lives at
 *    words[w * capacity + p].  Plain storage uses value planes
 *    vals[c * capacity + p] (float or double).
 *  - A handle (aq_sim) is not thread-safe; use one stream per handle.
 */
#ifndef AQ_MPM_H
#define AQ_MPM_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define AQ_OK 0
#define AQ_ERR_INVALID (-1)  /* bad argument or layout */
#define AQ_ERR_HIP (-2)      /* a HIP runtime call failed */
#define AQ_ERR_CAPACITY (-3) /* size beyond what the handle or workspace holds */

#define AQ_MAX_COMPS 32
#define AQ_MAX_WORDS 16

/* Quant type kinds (TS/python/taichi/lang/quant_impl.py:13-73,
   TS/taichi/ir/type.h:207-293). */
enum aq_kind {
  AQ_FIXED = 0, /* ti.quant.fixed: custom float without exponent, scale+offset */
  AQ_INT = 1,   /* ti.quant.int: custom int                                     */
  AQ_FLOAT = 2, /* ti.quant.float: exponent + digits (shared or not)            */
  AQ_F32 = 3    /* plain IEEE binary32 member (unquantized)                      */
};

/* One scalar member of a particle record after placement.
   Replaces the (CustomIntType/CustomFloatType, bit_offset, word_offset,
   exp_snode) tuple the reference keeps per place SNode
   (TS/taichi/ir/snode.h:121-145, TS/taichi/struct/struct_llvm.cpp:71-116). */
typedef struct aq_component {
  int32_t kind;       /* enum aq_kind */
  int32_t num_bits;   /* digit/int bits incl. sign; 32 for AQ_F32 */
  int32_t is_signed;
  int32_t bit_offset; /* absolute bit offset in the record, LSB first */
  int32_t exp_bits;   /* AQ_FLOAT: exponent width (<= 8) */
  int32_t exp_offset; /* AQ_FLOAT: absolute bit offset of the exponent field */
  int32_t shared;     /* AQ_FLOAT: exponent shared by all members with this exp_offset */
  int32_t use_dithering;
  double scale;       /* AQ_FIXED: range/2^(bits-1) signed, range/2^bits unsigned */
  double offset;      /* AQ_FIXED: offset */
} aq_component;

/* Container kinds of a placement (TS/python/taichi/lang/snode.py:108-180). */
enum aq_container_kind { AQ_BIT_STRUCT = 0, AQ_BIT_PACK = 1 };

/* One member to place.  `container` indexes the containers array;
   `shared_group` >= 0 puts the member in a shared-exponent placement group
   (SNode.place(..., shared_exponent=True)); members of one group must be
   consecutive in the places array and use the same exponent width. */
typedef struct aq_place {
  int32_t kind, num_bits, is_signed, exp_bits, use_dithering;
  int32_t container;
  int32_t shared_group;
  int32_t pad;
  double scale, offset;
} aq_place;

typedef struct aq_container {
  int32_t kind;     /* enum aq_container_kind */
  int32_t num_bits; /* AQ_BIT_STRUCT: 8, 16, 32 or 64; ignored for AQ_BIT_PACK */
} aq_container;

const char *aq_last_error(void);
/* Library version string and the GPU architecture it was built for. */
const char *aq_version(void);

/* Lay out places into a particle record of u64 words.  Follows
   struct_llvm.cpp:71-116 (offsets accumulate in placement order; bit_pack
   word = offset/64, members may straddle words; bit_struct members must fit
   the container) and snode_expr_utils.cpp:44-89 (a custom-float exponent is
   placed before its first user; a shared group has one exponent).
   Containers are aligned to their size (bit_pack and bit_struct(64): one
   word).  Unlike the reference, a bit_pack may hold custom floats with
   exponents (codegen_llvm_quant.cpp:766-767 forbids them there).
   out[i] describes places[i]; *nwords receives the record size. */
int aq_layout_plan(const aq_container *containers, int ncontainers,
                   const aq_place *places, int nplaces, aq_component *out,
                   int *nwords);

/* ---- codec: replaces the quantized load/store codegen ---------------------
   aq_pack: bit_pack_assign of whole records
   (TS/taichi/codegen/codegen_llvm_quant.cpp:884-952 via
   TS/python/taichi/lang/impl.py:753-756): vals[c * n + p] (device, f32)
   encoded with float_to_custom_int (:44-80) / shared exponents (:347-499)
   and written whole.  noise: NULL = counter-based noise of (seed, step,
   p, c); else device [ncomp][n] uniforms in [0, 1).
   aq_unpack: load_custom_float / load_custom_int (:759-784, :529-621). */
int aq_pack(const aq_component *comps, int ncomp, int nwords, const float *vals,
            const float *noise, uint64_t seed, uint32_t step, uint64_t *words,
            int64_t n, void *stream);
int aq_unpack(const aq_component *comps, int ncomp, int nwords,
              const uint64_t *words, float *vals, int64_t n, void *stream);
/* Masked store of the members selected by `mask` (bit c = component c) into
   records idx[k] (k < count); vals[j * count + k] is the value of the j-th
   selected member (increasing component order).  A non-shared member is a
   masked store (TS/taichi/codegen/codegen_llvm.cpp:1270-1300); a shared
   exponent group with a selected member is re-encoded, its unselected
   members reconstructed from the record first (codegen_llvm_quant.cpp:347-383).
   Selecting all members of one container is bit_pack_assign
   (TS/python/taichi/lang/impl.py:753-756).  idx, vals: device arrays;
   duplicate indices race. */
int aq_store_members(const aq_component *comps, int ncomp, int nwords,
                     uint64_t *words, int64_t n, uint32_t mask, const int64_t *idx,
                     const float *vals, int64_t count, uint64_t seed,
                     uint32_t step, void *stream);

/* ---- MLS-MPM simulation: replaces MPMSolver's kernels ---------------------
   (mpm_exps/engine/mpm_solver.py:19-122, 578-758). */
enum aq_material { AQ_FLUID = 0, AQ_ELASTIC = 1 };
enum aq_boundary { AQ_SEPARATING = 0, AQ_STICKY = 1 };
enum aq_storage { AQ_STORE_QUANT = 0, AQ_STORE_F32 = 1, AQ_STORE_F64 = 2 };

typedef struct aq_sim_desc {
  int32_t dim;      /* 2 or 3 */
  int32_t material; /* AQ_FLUID (J, sigma = la (J-1) I) or AQ_ELASTIC (2D fixed-corotated) */
  int32_t n_grid;   /* cells per axis; multiple of 8 (2D) or 4 (3D) */
  int32_t boundary; /* enum aq_boundary */
  int32_t bound;    /* wall thickness: lb = bound, ub = n_grid - bound */
  int32_t storage;  /* enum aq_storage */
  int32_t ncomp;    /* number of state components (2d + d^2 + 1 or 2d + 2d^2) */
  int32_t nwords;   /* AQ_STORE_QUANT: record size in u64 words */
  int64_t capacity; /* particle capacity = plane stride */
  double dx, inv_dx, dt, p_vol, p_mass, gravity; /* p_mass > 0 (AQ_ERR_INVALID otherwise) */
  double la;        /* fluid bulk modulus (MPMSolver.la) */
  double mu_0, la_0;/* elastic Lame parameters */
  uint64_t seed;    /* dithering seed */
  /* AQ_STORE_QUANT: components in the order of g2p_bls's bit_pack_assign
     (mpm_solver.py:700-744): x[d], v[d], C[d*d] row major, J or F[d*d]. */
  aq_component comps[AQ_MAX_COMPS];
  /* Per-axis walls (MPMConfig.lb / ub, mpm_solver.py:97-105): with
     axis_bounds = 1, axis a has walls at lb[a] and ub[a] (cells,
     0 <= lb[a] < ub[a] <= n_grid) instead of bound and n_grid - bound. */
  int32_t axis_bounds;
  int32_t lb[3], ub[3];
  /* use_bound_clip (mpm_solver.py:696-698): G2P clamps every new position
     component to [lb dx, ub dx]. */
  int32_t bound_clip;
  /* AQ_STORE_QUANT in 2D: compute in f64 (init_quantaichi's compute type,
     mpm_solver.py:372-374, f64 in the reference's 2D runs): fixed-point and
     int members are encoded and decoded in double, custom floats keep f32
     compute (as the reference's CustomFloatType); the grid is double.  The
     run-time specialised kernel is f32 only (aq_sim_specialize refuses). */
  int32_t compute_f64;
} aq_sim_desc;

typedef struct aq_sim aq_sim;

/* Workspace: grid node buffers, block flags, sort buffers, work items and
   one P2G tile per work item. */
int64_t aq_sim_workspace_bytes(const aq_sim_desc *desc);
int aq_sim_create(const aq_sim_desc *desc, void *workspace, int64_t workspace_bytes,
                  aq_sim **out);
int aq_sim_destroy(aq_sim *sim);

/* Records of particles [0, n) of a quantized simulation from values in its
   compute type (float, or double with compute_f64), planes vals[c * n + p],
   encoded as its kernels store them (dithering noise of `step`, keyed by
   the slot p); words[w * capacity + p].  aq_sim_unpack decodes them back.
   Replaces the Python-scope stores and loads of the reference's quantized
   fields (x[s, p] = ..., codegen_llvm_quant.cpp:884-952, 529-784) in the
   simulation's compute type; aq_pack / aq_unpack are the f32 codec of a
   bare layout. */
int aq_sim_pack(aq_sim *sim, const void *vals, int64_t n, uint32_t step, uint64_t *words, void *stream);
int aq_sim_unpack(aq_sim *sim, const uint64_t *words, int64_t n, void *vals, void *stream);

/* Run-time specialisation to the simulation's record layout (replaces the
   per-scheme kernel compilation of the reference's Taichi JIT,
   TS/taichi/codegen/codegen_llvm_quant.cpp: kernels are generated for each
   quantization scheme).  The fused G2P/P2G kernel of aq_advance and the
   P2G-only and G2P-only kernels that open and close a run of substeps (also
   those of aq_p2g / aq_g2p) are compiled with hipRTC from the kernel headers
   in src_dir (autoquantizer_amd/csrc) and include_dir (include/) with the
   layout's descriptors as constants, and used by every later tiled launch
   of `sim` (results identical to the generic kernels).  Code objects are cached in
   cache_dir (NULL: compile, no cache).  src_dir NULL returns to the generic
   kernel.  Quantized storage only. */
int aq_sim_specialize(aq_sim *sim, const char *src_dir, const char *include_dir, const char *cache_dir);
/* Compile (or find in cache_dir) the specialised kernel's code object for a
   descriptor without a device; path receives its file name. */
int aq_jit_build(const aq_sim_desc *desc, const char *src_dir, const char *include_dir, const char *cache_dir,
                 char *path, int64_t path_cap);

/* The reference substep (mpm_solver.py:624-631): build_pid (aq_sort),
   p2g_bls, grid_op, g2p_bls, as separate launches.  `state` is the words
   (AQ_STORE_QUANT) or value planes; n <= capacity.  `ids` (device
   int32[capacity] or NULL) are the particles' stable ids, permuted with the
   state by the sort; the dithering noise of a particle is keyed by its id
   (by its slot when ids is NULL) and by `step`.
   aq_p2g and aq_g2p use the LDS-tiled kernels when the last aq_sort on this
   state and particle count is still current, global-memory kernels
   otherwise; the results agree up to float rounding (accumulation order).
   After aq_p2g the grid (aq_sim_grid) holds (m v, m) per node; aq_grid_op
   turns it into velocities in place. */
int aq_substep(aq_sim *sim, void *state, int64_t n, int32_t *ids, uint32_t step, void *stream);
int aq_p2g(aq_sim *sim, const void *state, int64_t n, void *stream);
int aq_grid_op(aq_sim *sim, uint32_t step, void *stream);
int aq_g2p(aq_sim *sim, void *state, int64_t n, const int32_t *ids, uint32_t step, void *stream);

/* nsteps substeps with G2P of substep t fused with P2G of substep t+1
   (one launch reads the state once, writes it once and scatters the next
   grid).  Same results as nsteps aq_substep calls up to float rounding
   (accumulation order).  Particles are re-sorted by grid block at the
   call's start and then every `sort_every` substeps (0 = never:
   global-memory kernels), so every call with sort_every > 0 costs at least
   one sort: prefer few long calls to many short ones.  ids are permuted
   along. */
int aq_advance(aq_sim *sim, void *state, int64_t n, int nsteps, uint32_t step0,
               int sort_every, int32_t *ids, void *stream);

/* aq_advance with HIP events recorded on `stream` around every launch:
   ms[0] fused G2P->P2G kernels, ms[1] grid updates, ms[2] sorts, ms[3] the
   first P2G and last G2P (summed milliseconds); launches[0..3] the launch
   counts per category; *active_nodes the grid nodes in active blocks of the
   last grid.  Synchronises `stream`.  For measurement; the reference's
   counterpart is its kernel profiler (TS/taichi/program/kernel_profiler.h,
   TS/taichi/backends/cuda/cuda_profiler.cpp:10-44). */
int aq_advance_timed(aq_sim *sim, void *state, int64_t n, int nsteps, uint32_t step0,
                     int sort_every, int32_t *ids, float *ms, int32_t *launches,
                     int64_t *active_nodes, void *stream);

/* Sort particles by grid block of their stencil base (replaces build_pid's
   per-block lists, mpm_solver.py:578-588, node_dynamic.h:61-90); inside a
   block, in layers over its 64 cells (the r-th particle of every cell that
   has more than r, in cell order, then the (r+1)-th).  Permutes the state
   and ids in place through the workspace. */
int aq_sort(aq_sim *sim, void *state, int64_t n, int32_t *ids, void *stream);

/* Per-component max and min over particles, accumulated into
   mx_mn[0..ncomp) (max) and mx_mn[ncomp..2 ncomp) (min), device f32;
   replaces update_limits_* (mpm_solver.py:960-979). */
int aq_update_limits(aq_sim *sim, const void *state, int64_t n, float *mx_mn,
                     void *stream);
/* aq_advance with aq_update_limits after every substep (the range run of
   run_forward_detect_ranges, run_mpm_elastic_2d.py:174-184: detect_bound
   after each substep), in one call. */
int aq_advance_limits(aq_sim *sim, void *state, int64_t n, int nsteps, uint32_t step0, int sort_every,
                      int32_t *ids, float *mx_mn, void *stream);
/* compute_ek (mpm_solver.py:1043-1052) reduced in f64 into *out (device). */
int aq_energy(aq_sim *sim, const void *state, int64_t n, double *out, void *stream);

/* Range-violation counter (no counterpart in the reference, whose
   out-of-range fixed-point values wrap silently, codegen_llvm_quant.cpp:
   44-80): from now on every record encode of `sim` (G2P, fused G2P->P2G)
   adds to counts[c] (device, ncomp entries, zeroed by the caller) the
   stores of component c whose fixed-point or int code wrapped (the rounded
   value outside the member's range, or NaN).  counts NULL stops counting.
   Stores still wrap (bit-identical records either way). */
int aq_sim_count_wraps(aq_sim *sim, unsigned long long *counts);

/* ---- SDF collider of the 3D fluid scenes ----------------------------------
   Replaces init_sdf_collider / init_sdf_value (mpm_solver.py:284-319) and
   sdf_collision / sdf_collision_no_arg / rot_scene / step_collider /
   step_collider_difftaichi (:1444-1559).  A signed-distance volume (a
   rotating propeller in the reference) is placed in the grid; after the
   grid update of substep s (the `step` of aq_substep / aq_grid_op, step0 + t
   in aq_advance) every node I is mapped into the volume by rot_scene: the
   offset of I from the rotation centre c (x, z) is rotated by -theta, the
   height subtracted, the indices truncated and clamped to the walls.  Where
   the volume's value is below -1e-5 the node is inside: with normal n (the
   stored gradient rotated by theta) and collider velocity omega e_y x
   ((I - c) dx, 0 in y) (or of the rotated index when
   velocity_from_rotated), the velocity relative to the collider loses its
   component into the collider: v = v_c + v_rel - min(v_rel . n, 0) n.
   With s counted from first_step, the collider stands still during
   substep 0 (collider_omega is set by its first step) and has taken
   k = max(s - lag, 0) steps at substep s:
   theta = omega dt k, height = height_rate dt k (lag 1, rate 0.15: the
   range and AD runs' step_collider_difftaichi; lag 0, rate 0.1: the
   forward runs' step_collider).  3D only; the arrays are device memory the
   caller keeps alive while the collider is set; NULL removes it. */
typedef struct aq_collider {
  const float *sdf;     /* [sx][sy][sz] signed distance (collider_sdf) */
  const float *normal;  /* [sx][sy][sz][3] its gradient (collider_sdf_normal) */
  int32_t size[3];      /* sx, sy, sz */
  int32_t offset[3];    /* node offset + (i, j, k) holds entry (i, j, k) (sdf_offset) */
  int32_t center[3];    /* rotation centre in nodes (get_center_idx; y unused) */
  int32_t lag;
  int32_t velocity_from_rotated;
  uint32_t first_step;  /* the step of the collider's substep 0 */
  double omega;         /* angular velocity about +y */
  double height_rate;
} aq_collider;
int aq_sim_set_collider(aq_sim *sim, const aq_collider *collider);

/* ---- groups: many quantized simulations advanced together ---------------
   Replaces the reference's one-process-per-candidate evaluation of bit
   allocations (mpm_exps/run_2d_exp.py:46-70 starting
   run_mpm_optimality_check.py per candidate).  A group holds simulations
   with quantized storage (any record layouts, dimensions and materials),
   each with its own state, particle count and ids, fixed at creation.  Each
   substep's fused G2P->P2G runs as one launch per class of simulations
   (dimension, material, record-word bucket) whose workgroups walk the
   items of all the class's simulations, and the gather + grid update as
   one launch per dimension over all their grid blocks; sorts, the
   first P2G and the last G2P run per simulation.  Results equal
   aq_advance's on each simulation with the same sort interval and step0
   up to float rounding (accumulation order; the untiled fallback of edge
   particles).  The group reads its simulations' settings at creation (the
   range-violation counter included): create it after aq_sim_count_wraps.
   The workspace (device, caller-owned) holds the launch descriptors. */
typedef struct aq_group aq_group;
int64_t aq_group_workspace_bytes(int nsims);
int aq_group_create(aq_sim *const *sims, void *const *states, const int64_t *ns, int32_t *const *ids, int nsims,
                    void *workspace, int64_t workspace_bytes, aq_group **out);
/* nsteps substeps of every simulation (dithering step of substep t:
   step0 + t), sorted every `sort_every` substeps (> 0); `workgroups` is the
   grid of the fused launches (a multiple of 8 up to 8192; 0: sized from the
   particle count). */
int aq_group_advance(aq_group *g, int nsteps, uint32_t step0, int sort_every, int workgroups, void *stream);
int aq_group_destroy(aq_group *g);

/* Device pointer of the grid node buffer (float4 or double4 per node,
   blocked layout: (m v, m) after aq_p2g, (v, m) after the grid update;
   zero outside the blocks the last P2G touched) and its node count.  The
   next G2P reads this buffer. */
int aq_sim_grid(aq_sim *sim, void **grid, int64_t *nodes);

/* ---- adjoint of the plain substep: replaces the reference's Taichi-autodiff
   p2g.grad / grid_op.grad / sdf_collision_no_arg.grad / g2p.grad and
   accu_grad (mpm_solver.py:800-858, 1167-1198, 1232-1271, 1476-1504) ----- */
/* Workspace of aq_adj_substep: the P2G grid and the grid adjoint. */
int64_t aq_adj_workspace_bytes(const aq_sim *sim);
/* One reverse substep on plain storage (AQ_STORE_F32 or AQ_STORE_F64):
   adj_out = dL/ds_t from adj_next = dL/ds_{t+1}, for the state s_t in
   `state` (value planes, stride capacity; adj_next and adj_out alike).
   Recomputes s_t's P2G (LDS-tiled when the last aq_sort of `sim` sorted
   this buffer, else untiled) and grid update (with the SDF collider of
   substep `step`, the step aq_advance ran s_t -> s_{t+1} at), then runs the
   G2P, grid (walls and collider) and P2G adjoint kernels.  If gsum (device,
   double[ncomp]) is not NULL, sum_p adj_out[c][p]^2 is added to gsum[c]
   (the reference's accu_grad).  The bound clip has no adjoint here
   (AQ_ERR_INVALID). */
int aq_adj_substep(aq_sim *sim, uint32_t step, const void *state, int64_t n, const void *adj_next,
                   void *adj_out, double *gsum, void *workspace, int64_t workspace_bytes, void *stream);

/* Modelled quantization noise of the adjoint error-gradient run: each
   quantized store after a substep adds a uniform rounding error of width
   sigma[c] (the model behind the reference's error sums,
   mpm_solver.py:1232-1271 and solvers/analytical_solver.py:3-84, applied
   here as explicit noise so that d loss / d sigma can be taken and checked
   by finite differences).  xi(step, c, p) is uniform in [-1/2, 1/2), the
   dithering noise of (seed, step, component c, particle id) minus 1/2; the
   id is ids[slot] (device int32, the ids aq_sort permutes with the state)
   or the slot when ids is NULL, so a sorted state draws the same noise per
   particle as the unsorted one.
   aq_noise_add: state[c][p] += sigma[c] xi (sigma: host double[ncomp],
   finite, >= 0).  aq_noise_dot: out[c] += sum_p adj[c][p] xi (out: device
   double[ncomp]), the derivative of a loss w.r.t. sigma[c] for the noise
   added at `step` when adj = dL/d(state after the noise).  Plain storage. */
int aq_noise_add(aq_sim *sim, void *state, int64_t n, const int32_t *ids, const double *sigma, uint64_t seed,
                 uint32_t step, void *stream);
int aq_noise_dot(aq_sim *sim, const void *adj, int64_t n, const int32_t *ids, uint64_t seed, uint32_t step,
                 double *out, void *stream);

/* Codec self-check (no counterpart in the reference): the fused kernels'
   encoder decodes the shared-exponent digits it writes with a shorter
   formula than the record decode (reconstruct_custom_float_with_exponent,
   TS/taichi/codegen/codegen_llvm_quant.cpp:660-757); adds to *mismatches
   (device u64, zero it first) the digit patterns, over every exponent field
   and digit pattern of signed members with 4..8 exponent bits and 2..20
   digit bits, on which the two differ in any bit. */
int aq_codec_selftest(unsigned long long *mismatches, void *stream);

/* ---- surrogate MLP of the scheme search: training and inference kernels
   (no counterpart in the reference, whose search always runs the adjoint) ---
   in -> 64 (tanh) -> 64 (tanh) -> 64 (tanh) -> out; parameters one flat fp32
   array W1[64][in], b1, W2[64][64], b2, W3[64][64], b3, W4[out][64], b4. */
#define AQ_MLP_IN 24
#define AQ_MLP_HIDDEN 64
#define AQ_MLP_OUT 4
#define AQ_MLP_PARAMS 10180
/* y[s][out] for x[s][in]; act [s][3 * 64] receives the hidden activations
   (h1, h2, h3) that aq_mlp_backward uses. */
int aq_mlp_forward(const float *params, const float *x, int32_t nsamples, float *act, float *y, void *stream);
/* Gradient of mean((y - target)^2) over samples and outputs plus
   weight_decay * (sum of the squared weight-matrix entries; biases excluded)
   with respect to the parameters, into grad[AQ_MLP_PARAMS]; delta: scratch
   [s][3 * 64 + out]; loss (NULL or device double): the summed squared error
   (without the decay term) is added to it. */
int aq_mlp_backward(const float *params, const float *x, const float *act, const float *y,
                    const float *target, int32_t nsamples, float weight_decay, float *delta, float *grad,
                    double *loss, void *stream);
/* aq_mlp_backward without the delta scratch argument, for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops,
   harnesses that forbid hipMalloc).  No counterpart in the reference.  The
   deltas live in a static device array of AQ_MLP_NOALLOC_MAX_SAMPLES
   samples (1.6 MB per GPU, held until the process ends); nsamples above
   that returns AQ_ERR_CAPACITY.  Allocates nothing.  All calls on one GPU
   share the array: calls that overlap in time on one GPU (other streams or
   threads) race.  Results equal aq_mlp_backward's. */
#define AQ_MLP_NOALLOC_MAX_SAMPLES 2048
int aq_mlp_backward_noalloc(const float *params, const float *x, const float *act, const float *y,
                            const float *target, int32_t nsamples, float weight_decay, float *grad,
                            double *loss, void *stream);
/* One Adam step (bias-corrected, step t >= 1) on n parameters. */
int aq_adam_step(float *params, const float *grad, float *m, float *v, int32_t n, int32_t t, float lr,
                 float beta1, float beta2, float eps, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* AQ_MPM_H */
