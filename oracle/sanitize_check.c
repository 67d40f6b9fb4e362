/*
 * Sanitizer run of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc, host
 * code only; `make -C oracle sanitize`) and run by
 * tests/test_oracle_sanitizers.py: every entry point the parity tests use
 * is exercised on small inputs, so an out-of-bounds access, a shift past a
 * word, a signed overflow or a misaligned load in the oracle stops the run.
 * It also checks two codec properties the reference's bit_pack relies on
 * (mpm_exps/engine/mpm_solver.py:700-744, TS/taichi/codegen/codegen_llvm_quant.cpp:920-950):
 *   - a packed store of a particle range leaves every other particle's
 *     words untouched, and a record whose size is a multiple of 64 bits
 *     writes no word past it (the arrays are allocated exactly, so ASan
 *     sees a stray word);
 *   - decode is a projection: encoding decoded values and decoding again
 *     gives the same values (NaN, Inf and out-of-range inputs included),
 *     up to one step for dithered members.
 * Prints "sanitize check ok" and exits 0 when everything passed.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_mpm.h"

static int failures = 0;
#define CHECK(cond, ...)                              \
  do {                                                \
    if (!(cond)) {                                    \
      fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                   \
      fprintf(stderr, "\n");                          \
      failures++;                                     \
    }                                                 \
  } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand(void) { /* splitmix64 -> [0, 1) */
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

static orc_comp fixed(int nbits, int is_signed, int off, double range, double offset, int dither) {
  orc_comp c;
  memset(&c, 0, sizeof c);
  c.kind = ORC_FIXED;
  c.nbits = nbits;
  c.is_signed = is_signed;
  c.bit_offset = off;
  c.dither = dither;
  c.scale = range / ldexp(1.0, is_signed ? nbits - 1 : nbits);
  c.offset = offset;
  return c;
}

static orc_comp custom_int(int nbits, int is_signed, int off) {
  orc_comp c;
  memset(&c, 0, sizeof c);
  c.kind = ORC_INT;
  c.nbits = nbits;
  c.is_signed = is_signed;
  c.bit_offset = off;
  c.scale = 1.0;
  return c;
}

static orc_comp custom_float(int nbits, int off, int exp_bits, int exp_off, int shared) {
  orc_comp c;
  memset(&c, 0, sizeof c);
  c.kind = ORC_FLOAT;
  c.nbits = nbits;
  c.is_signed = 1;
  c.bit_offset = off;
  c.exp_bits = exp_bits;
  c.exp_offset = exp_off;
  c.shared = shared;
  c.scale = 1.0;
  return c;
}

static orc_comp f32_at(int off) {
  orc_comp c;
  memset(&c, 0, sizeof c);
  c.kind = ORC_F32;
  c.nbits = 32;
  c.is_signed = 1;
  c.bit_offset = off;
  c.scale = 1.0;
  return c;
}

/* A two-word record with every member kind: fixed point (signed dithered,
   unsigned with an offset), custom int, custom float with its own exponent,
   two custom floats sharing one, and an fp32 member ending at bit 128. */
static int codec_layout(orc_comp *c) {
  int k = 0;
  c[k++] = fixed(20, 1, 0, 2.0, 0.0, 1);
  c[k++] = fixed(12, 0, 20, 1.0, 0.25, 0);
  c[k++] = custom_int(9, 1, 32);
  c[k++] = custom_float(10, 41, 5, 51, 0);
  c[k++] = custom_float(12, 64, 6, 56, 1);
  c[k++] = custom_float(12, 76, 6, 56, 1);
  c[k++] = f32_at(96);
  return k;
}

static void check_codec(void) {
  enum { N = 777, NW = 2 };
  orc_comp comps[8];
  const int nc = codec_layout(comps);
  float *vals = xcalloc((size_t)nc * N, sizeof(float));
  float *v1 = xcalloc((size_t)nc * N, sizeof(float));
  float *v2 = xcalloc((size_t)nc * N, sizeof(float));
  uint64_t *w = xcalloc((size_t)NW * N, sizeof(uint64_t));
  uint64_t *w2 = xcalloc((size_t)NW * N, sizeof(uint64_t));
  const float specials[] = {NAN, INFINITY, -INFINITY, 1e30f, -1e30f, 0.0f, -0.0f, 1e-40f};
  for (int c = 0; c < nc; c++)
    for (int p = 0; p < N; p++) {
      float x = (float)(4.0 * urand() - 2.0);
      if (p < 8) x = specials[p];
      if (p >= 8 && p < 16) x *= 1000.0f; /* out of every fixed range */
      vals[(size_t)c * N + p] = x;
    }
  orc_pack_f32(comps, nc, NW, vals, NULL, 7u, 3u, w, N);
  orc_unpack_f32(comps, nc, NW, w, v1, N);
  orc_pack_f32(comps, nc, NW, v1, NULL, 7u, 4u, w2, N);
  orc_unpack_f32(comps, nc, NW, w2, v2, N);
  /* dithered members may move by one step: x / scale + u is rounded to f32
     before the floor, so u close to 1 can carry an exact code up */
  int diff = 0, far = 0;
  for (int c = 0; c < nc; c++)
    for (int p = 0; p < N; p++) {
      const size_t i = (size_t)c * N + p;
      if (comps[c].kind == ORC_FIXED && comps[c].dither)
        far += !(fabs((double)v2[i] - (double)v1[i]) <= comps[c].scale * 1.0000001);
      else
        diff += memcmp(&v1[i], &v2[i], sizeof(float)) != 0;
    }
  CHECK(diff == 0, "decode is not a projection: %d values changed on re-encoding", diff);
  CHECK(far == 0, "%d dithered values moved by more than one step on re-encoding", far);

  /* a range store leaves the other particles' words as they were */
  memcpy(w2, w, sizeof(uint64_t) * NW * N);
  const int64_t lo = 100, hi = 357;
  for (int64_t i = 0; i < (int64_t)nc * N; i++) vals[i] = -vals[i];
  orc_pack_range_f32(comps, nc, NW, vals, NULL, 7u, 5u, w2, N, lo, hi);
  int outside = 0, inside = 0;
  for (int wd = 0; wd < NW; wd++)
    for (int64_t p = 0; p < N; p++) {
      const int same = w2[(size_t)wd * N + p] == w[(size_t)wd * N + p];
      if (p < lo || p >= hi)
        outside += !same;
      else
        inside += !same;
    }
  CHECK(outside == 0, "range store changed %d words outside [%lld, %lld)", outside, (long long)lo, (long long)hi);
  CHECK(inside > 0, "range store changed nothing inside its range");
  orc_unpack_range_f32(comps, nc, NW, w2, v2, N, lo, hi);

  /* single-member stores into one record */
  uint64_t rec[NW] = {0, 0};
  for (int c = 0; c < nc; c++) orc_store_one_f32(comps, nc, rec, c, (float)(0.3 * (c + 1)), 0.5f);

  /* the same layout in f64 */
  double *dv = xcalloc((size_t)nc * N, sizeof(double));
  double *dv1 = xcalloc((size_t)nc * N, sizeof(double));
  for (int64_t i = 0; i < (int64_t)nc * N; i++) dv[i] = (double)v1[i];
  orc_pack_f64(comps, nc, NW, dv, NULL, 9u, 1u, w2, N);
  orc_unpack_f64(comps, nc, NW, w2, dv1, N);
  for (int64_t i = 0; i < (int64_t)nc * N; i++) CHECK(!isnan(dv1[i]) || isnan(dv[i]), "f64 decode made a NaN");

  /* scalar codec entry points at their edges */
  for (int nb = 1; nb <= 32; nb++) {
    const double sc = 1.0 / ldexp(1.0, nb - 1);
    for (int sg = 0; sg < 2; sg++) {
      const uint32_t code = orc_fixed_encode_f64(0.75, sc, 0.0, nb, sg, 1, 0.5);
      (void)orc_fixed_decode_f64(code, nb, sg, sc, 0.0);
      (void)orc_fixed_decode_f32(orc_fixed_encode_f32(-1e9f, sc, 0.0, nb, sg, 0, 0.0f), nb, sg, sc, 0.0);
      (void)orc_sext(code, nb);
    }
  }
  for (int nb = 2; nb <= 24; nb++)
    for (int eb = 1; eb <= 8; eb++) {
      uint32_t e = 0, dg = 0;
      const float xs[4] = {0.0f, 1.5f, -3.25e-3f, 7e20f};
      for (int i = 0; i < 4; i++) {
        orc_float_encode(xs[i], nb, 1, eb, &e, &dg);
        (void)orc_float_decode(e, dg, nb, 1, eb);
      }
      const int32_t nbs[3] = {nb, nb, nb}, sgs[3] = {1, 1, 1};
      uint32_t ds[3];
      orc_shared_encode(xs, 3, nbs, sgs, eb, &e, ds);
      for (int i = 0; i < 3; i++) (void)orc_shared_decode(e, ds[i], nb, 1, eb);
    }
  for (int off = 0; off + 32 <= 128; off += 7) {
    uint64_t r2[2] = {0, 0};
    orc_put_bits(r2, off, 32, 0xDEADBEEFu);
    CHECK(orc_get_bits(r2, off, 32) == 0xDEADBEEFu, "bit field at %d does not round trip", off);
  }
  free(vals), free(v1), free(v2), free(w), free(w2), free(dv), free(dv1);
}

static orc_mpm scene(int dim, int material, int ng) {
  orc_mpm s;
  memset(&s, 0, sizeof s);
  s.dim = dim;
  s.material = material;
  s.n_grid = ng;
  s.boundary = ORC_SEPARATING;
  s.bound = 3;
  s.dx = 1.0 / ng;
  s.inv_dx = ng;
  s.dt = 1e-4;
  s.p_vol = pow(s.dx * 0.5, dim);
  s.p_mass = s.p_vol * 1.0;
  s.gravity = 9.8;
  s.la = 10.0;
  s.mu_0 = 1e3 / (2 * 1.2);
  s.la_0 = 1e3 * 0.2 / (1.2 * 0.6);
  return s;
}

/* particles uniform in [0.3, 0.6]^dim with small velocities, C = 0 and the
   rest state (J = 1 or F = I); component-major */
static void fill_state(const orc_mpm *s, double *st, int64_t n) {
  const int d = s->dim, nc = orc_num_comps(s);
  memset(st, 0, sizeof(double) * (size_t)nc * (size_t)n);
  for (int64_t p = 0; p < n; p++) {
    for (int a = 0; a < d; a++) st[(int64_t)a * n + p] = 0.3 + 0.3 * urand();
    for (int a = 0; a < d; a++) st[(int64_t)(d + a) * n + p] = 0.2 * urand() - 0.1;
    const int base = 2 * d + d * d;
    if (s->material == ORC_FLUID)
      st[(int64_t)base * n + p] = 1.0;
    else
      for (int a = 0; a < d; a++) st[(int64_t)(base + a * d + a) * n + p] = 1.0;
  }
}

static size_t grid_nodes(const orc_mpm *s) {
  const size_t g = (size_t)s->n_grid;
  return s->dim == 3 ? g * g * g : g * g;
}

/* a ball-shaped collider volume of 8^3 entries near the middle of the domain */
typedef struct {
  float sdf[8 * 8 * 8];
  float normal[8 * 8 * 8 * 3];
} ball;

static orc_collider make_collider(const orc_mpm *s, ball *b, double theta) {
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++)
      for (int k = 0; k < 8; k++) {
        const int e = (i * 8 + j) * 8 + k;
        const double r[3] = {i - 3.5, j - 3.5, k - 3.5};
        const double len = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        b->sdf[e] = (float)((len - 3.0) * s->dx);
        for (int a = 0; a < 3; a++) b->normal[3 * e + a] = (float)(r[a] / (len > 0 ? len : 1.0));
      }
  orc_collider c;
  memset(&c, 0, sizeof c);
  c.sdf = b->sdf;
  c.normal = b->normal;
  for (int a = 0; a < 3; a++) {
    c.size[a] = 8;
    c.offset[a] = s->n_grid / 2 - 4;
    c.center[a] = s->n_grid / 2;
  }
  c.from_rot = 1;
  c.yshift = 1;
  c.cos_t = cos(theta);
  c.sin_t = sin(theta);
  c.omega = 2.0;
  return c;
}

static void check_mpm(int dim, int material, int ng, int64_t n) {
  orc_mpm s = scene(dim, material, ng);
  const int nc = orc_num_comps(&s);
  const size_t nodes = grid_nodes(&s);
  double *st = xcalloc((size_t)nc * (size_t)n, sizeof(double));
  fill_state(&s, st, n);
  float *sf = xcalloc((size_t)nc * (size_t)n, sizeof(float));
  for (int64_t i = 0; i < (int64_t)nc * n; i++) sf[i] = (float)st[i];
  float *gm = xcalloc(nodes, sizeof(float)), *gv = xcalloc(nodes * (size_t)dim, sizeof(float));
  double *gmd = xcalloc(nodes, sizeof(double)), *gvd = xcalloc(nodes * (size_t)dim, sizeof(double));

  /* the phases one by one, then whole substeps */
  float *out = xcalloc((size_t)nc * (size_t)n, sizeof(float));
  orc_p2g_f32(&s, sf, n, gm, gv);
  orc_grid_op_f32(&s, gm, gv);
  orc_g2p_f32(&s, sf, gv, out, n);
  double *outd = xcalloc((size_t)nc * (size_t)n, sizeof(double));
  orc_p2g_f64(&s, st, n, gmd, gvd);
  orc_grid_op_f64(&s, gmd, gvd);
  orc_g2p_f64(&s, st, gvd, outd, n);
  orc_run_plain_f32(&s, sf, n, 4, gm, gv, 2);
  orc_run_plain_f64(&s, st, n, 4, gmd, gvd, 2);
  const double e32 = orc_energy_f32(&s, sf, n), e64 = orc_energy_f64(&s, st, n);
  CHECK(isfinite(e32) && isfinite(e64), "energy not finite (%d-D material %d)", dim, material);
  float *mx = xcalloc((size_t)nc, sizeof(float)), *mn = xcalloc((size_t)nc, sizeof(float));
  orc_limits_f32(&s, sf, n, mx, mn);
  for (int c = 0; c < nc; c++) CHECK(mn[c] <= mx[c], "limits of component %d out of order", c);

  /* quantized storage: fixed-point x and v, fp32 for the rest */
  orc_comp *comps = xcalloc((size_t)nc, sizeof(orc_comp));
  int off = 0;
  for (int c = 0; c < nc; c++) {
    if (c < dim)
      comps[c] = fixed(16, 0, off, 1.0, 0.0, 1);
    else if (c < 2 * dim)
      comps[c] = fixed(16, 1, off, 4.0, 0.0, 1);
    else
      comps[c] = f32_at(off);
    off += comps[c].nbits;
  }
  const int nw = (off + 63) / 64;
  uint64_t *words = xcalloc((size_t)nw * (size_t)n, sizeof(uint64_t));
  orc_pack_f32(comps, nc, nw, sf, NULL, 11u, 0u, words, n);
  float *scratch = xcalloc((size_t)2 * nc * (size_t)n, sizeof(float));
  orc_run_quant_f32(&s, comps, nw, words, n, 3, 11u, 1u, gm, gv, scratch, 1);
  double *scratchd = xcalloc((size_t)2 * nc * (size_t)n, sizeof(double));
  orc_run_quant_f64(&s, comps, nw, words, n, 3, 11u, 4u, gmd, gvd, scratchd, 1);
  orc_unpack_f32(comps, nc, nw, words, out, n);
  int bad = 0;
  for (int64_t i = 0; i < (int64_t)nc * n; i++) bad += !isfinite(out[i]);
  CHECK(bad == 0, "%d non-finite values after quantized substeps (%d-D)", bad, dim);

  /* adjoint: one substep, then a short run */
  fill_state(&s, st, n);
  double *adj_next = xcalloc((size_t)nc * (size_t)n, sizeof(double));
  double *adj = xcalloc((size_t)nc * (size_t)n, sizeof(double));
  for (int64_t i = 0; i < (int64_t)nc * n; i++) adj_next[i] = urand() - 0.5;
  double *work = xcalloc((size_t)orc_adj_work_doubles(&s), sizeof(double));
  orc_adj_substep_f64(&s, st, n, adj_next, adj, work);
  double *gsum = xcalloc((size_t)nc, sizeof(double));
  const double L = orc_adjoint_run_f64(&s, st, n, 3, 0, NULL, adj, gsum);
  CHECK(isfinite(L), "adjoint run loss not finite");
  const double L1 = orc_adjoint_run_f64(&s, st, n, 2, 1, adj_next, adj, gsum);
  CHECK(isfinite(L1), "adjoint run (linear loss) not finite");

  if (dim == 3) {
    ball *b = xcalloc(1, sizeof(ball));
    const orc_collider col = make_collider(&s, b, 0.4);
    orc_collide_f32(&s, &col, gv);
    orc_collide_f64(&s, &col, gvd);
    orc_adj_substep_col_f64(&s, &col, st, n, adj_next, adj, work);
    int I[3] = {ng / 2, ng / 2 + 1, ng / 2}, hits = 0;
    float nf[3], cf[3];
    double nd[3], cd[3];
    for (int i = 0; i < ng; i++) {
      I[0] = i;
      hits += orc_collider_probe_f32(&s, &col, I, nf, cf);
      (void)orc_collider_probe_f64(&s, &col, I, nd, cd);
    }
    CHECK(hits > 0, "the collider covers no node on its centre line");
    free(b);
  }
  free(st), free(sf), free(gm), free(gv), free(gmd), free(gvd), free(out), free(outd), free(mx), free(mn);
  free(comps), free(words), free(scratch), free(scratchd), free(adj_next), free(adj), free(work), free(gsum);
}

int main(int argc, char **argv) {
  if (argc > 1 && strcmp(argv[1], "--plant-overflow") == 0) {
    /* a deliberate heap overflow: shows that the sanitizers are live */
    orc_comp comps[8];
    const int nc = codec_layout(comps);
    float vals[8] = {0};
    uint64_t *w = xcalloc(1, sizeof(uint64_t)); /* one word for a two-word record */
    orc_pack_f32(comps, nc, 2, vals, NULL, 1u, 0u, w, 1);
    free(w);
    return 0;
  }
  check_codec();
  check_mpm(2, ORC_ELASTIC, 32, 600);
  check_mpm(2, ORC_FLUID, 32, 600);
  check_mpm(3, ORC_FLUID, 16, 500);
  check_mpm(3, ORC_ELASTIC, 16, 500);
  /* noise helpers */
  double xi[3 * 64];
  orc_noise_fill(5u, 2u, 3, 64, xi);
  for (int i = 0; i < 3 * 64; i++) CHECK(xi[i] >= -0.5 && xi[i] < 0.5, "noise out of range");
  (void)orc_noise_u(5u, 2u, 7u, 1u);
  (void)orc_dither_u(5u, 2u, 7u, 1u);
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("sanitize check ok\n");
  return 0;
}
