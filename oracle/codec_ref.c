/*
 * Oracle codec: a plain-C restatement of the reference's quantized loads and
 * stores.  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Reference sources followed (TS/ = taichi_submission/):
 *   TS/taichi/codegen/codegen_llvm_quant.cpp   encode/decode/pack/split
 *   TS/taichi/runtime/llvm/runtime.cpp:1732-1762 rounding + dithering
 *   TS/taichi/ir/type.cpp:114-176               type parameters
 *   TS/python/taichi/lang/quant_impl.py:13-73   scale of ti.quant.fixed
 *   TS/taichi/struct/struct_llvm.cpp:71-116     bit_pack placement
 * Build with -ffp-contract=off: every float operation below is rounded
 * separately, as in the reference's unfused semantics.
 */
#include "orc.h"

#include <math.h>
#include <string.h>

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t mask32(int nb) { return nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u); }
static uint64_t mask64(int nb) { return nb >= 64 ? ~0ull : ((1ull << nb) - 1ull); }

int32_t orc_sext(uint32_t code, int nbits) {
  /* extract_custom_int_by_mask, codegen_llvm_quant.cpp:582-621: the sign bit
     nbits-1 is replicated upward */
  if (nbits >= 32) return (int32_t)code;
  code &= mask32(nbits);
  if (code & (1u << (nbits - 1))) code |= ~mask32(nbits);
  return (int32_t)code;
}

/* fptosi / fptoui (codegen_llvm_quant.cpp:74-78).  Inside the int range this
   is truncation toward zero; outside it (LLVM poison) we saturate and map NaN
   to 0. */
static int32_t cvt_i32_f64(double y) {
  if (y != y) return 0;
  if (y >= 2147483648.0) return 2147483647;
  if (y <= -2147483649.0) return (int32_t)0x80000000u;
  return (int32_t)y;
}
static uint32_t cvt_u32_f64(double y) {
  if (!(y > 0.0)) return 0u;
  if (y >= 4294967296.0) return 0xFFFFFFFFu;
  return (uint32_t)y;
}

/* rounding_prepare_f32 (runtime.cpp:1732-1742): add 0.5 carrying the sign
   bit of the argument.  The _with_dithering variant (runtime.cpp:1755-1762)
   first does f = f + rand - 0.5, evaluated left to right. */
static float round_prep_f32(float f, int dither, float u) {
  if (dither) {
    f = f + u;
    f = f - 0.5f;
  }
  uint32_t delta = (f2u(f) & 0x80000000u) | f2u(0.5f);
  return f + u2f(delta);
}
static double round_prep_f64(double f, int dither, double u) {
  if (dither) {
    f = f + u;
    f = f - 0.5;
  }
  return f + copysign(0.5, f);
}

/* float_to_custom_int (codegen_llvm_quant.cpp:44-80) followed by the
   num_bits mask of custom_type_to_bits (:175-181). */
uint32_t orc_fixed_encode_f32(float v, double scale, double offset, int nbits,
                              int is_signed, int dither, float u) {
  float s = (float)(1.0 / scale); /* :50-54, 1/scale in double then cast */
  float o = (float)offset;        /* :55-56 */
  float centered = v - o;         /* :57 */
  float scaled = centered * s;    /* :58 */
  scaled = round_prep_f32(scaled, dither, u);
  uint32_t q = is_signed ? (uint32_t)cvt_i32_f64((double)scaled)
                         : cvt_u32_f64((double)scaled);
  return q & mask32(nbits);
}

uint32_t orc_fixed_encode_f64(double v, double scale, double offset, int nbits,
                              int is_signed, int dither, double u) {
  double s = 1.0 / scale;
  double o = offset;
  double centered = v - o;
  double scaled = centered * s;
  scaled = round_prep_f64(scaled, dither, u);
  uint32_t q = is_signed ? (uint32_t)cvt_i32_f64(scaled) : cvt_u32_f64(scaled);
  return q & mask32(nbits);
}

/* reconstruct_custom_float (codegen_llvm_quant.cpp:624-641):
   (T)q * (T)scale + (T)offset, two rounded operations. */
float orc_fixed_decode_f32(uint32_t code, int nbits, int is_signed,
                           double scale, double offset) {
  float c = is_signed ? (float)orc_sext(code, nbits)
                      : (float)(code & mask32(nbits));
  float s = (float)scale;
  float t = c * s;
  float o = (float)offset;
  return t + o;
}

double orc_fixed_decode_f64(uint32_t code, int nbits, int is_signed,
                            double scale, double offset) {
  double c = is_signed ? (double)orc_sext(code, nbits)
                       : (double)(code & mask32(nbits));
  double t = c * scale;
  return t + offset;
}

/* CustomFloatType::get_exponent_conversion_offset (type.cpp:167-171) */
static int32_t exp_conv(int exp_bits) { return 127 - (1 << (exp_bits - 1)) + 1; }

/* Non-shared custom float store (codegen_llvm_quant.cpp:232-301). */
void orc_float_encode(float f, int nbits, int is_signed, int exp_bits,
                      uint32_t *exp_out, uint32_t *digits_out) {
  int db = nbits - is_signed; /* type.cpp:173-176 */
  uint32_t b = f2u(f);
  if (db < 23) b += 1u << (22 - db); /* :243-246 round to nearest */
  uint32_t e = (uint32_t)((int32_t)b >> 23) & 0xFFu;                  /* :248-250 */
  uint32_t d = (uint32_t)((int32_t)b >> (23 - db)) & mask32(db);       /* :251-255 */
  if (is_signed) d |= (b & 0x80000000u) >> (31 - db);                  /* :257-264 */
  int32_t es = (int32_t)e - (e != 0 ? exp_conv(exp_bits) : 0);        /* :269-271 */
  if (es < 0) es = 0;                                                  /* :272-273 */
  if (es == 0) d = 0;                                                  /* :289-297 FTZ */
  *exp_out = (uint32_t)es & mask32(exp_bits);
  *digits_out = d & mask32(nbits);
}

/* reconstruct_custom_float_with_exponent, shared_exponent=false
   (codegen_llvm_quant.cpp:660-757). */
float orc_float_decode(uint32_t exp, uint32_t digits, int nbits, int is_signed,
                       int exp_bits) {
  int db = nbits - is_signed;
  uint32_t off = (uint32_t)exp_conv(exp_bits);
  if (exp_bits < 8 && exp == 0) off = 0;     /* :675-681 */
  uint32_t d = digits & mask32(nbits);       /* :688-691 */
  d <<= (23 - db);                           /* :725-727 */
  uint32_t frac = d & 0x7FFFFFu;             /* :729 */
  uint32_t ev = exp + off;                   /* :731 */
  uint32_t fb = (ev << 23) | frac;           /* :733-736 */
  if (is_signed) fb |= (d & (1u << 23)) << 8; /* :749-754 */
  return u2f(fb);
}

/* store_floats_with_shared_exponents + get_float_digits_with_shared_exponents
   (codegen_llvm_quant.cpp:347-457, 476-499). */
void orc_shared_encode(const float *f, int n, const int32_t *nbits,
                       const int32_t *is_signed, int exp_bits,
                       uint32_t *exp_out, uint32_t *digits_out) {
  uint32_t m = 0;
  for (int k = 0; k < n; k++) {               /* :384-392 max_u32 of exponents */
    uint32_t e = (f2u(f[k]) >> 23) & 0xFFu;
    if (e > m) m = e;
  }
  int32_t es = (int32_t)m - (m != 0 ? exp_conv(exp_bits) : 0); /* :394-401 */
  if (es < 0) es = 0;
  *exp_out = (uint32_t)es & mask32(exp_bits);
  for (int k = 0; k < n; k++) {
    uint32_t b = f2u(f[k]);
    uint32_t e = (b >> 23) & 0xFFu;
    uint32_t sh = m - e;                       /* :480 */
    uint32_t d = (b & 0x7FFFFFu) | (1u << 23); /* :459-468 full digits, and :490-494 */
    if (sh > 31) sh = 31;                      /* :495 min_u32 */
    d >>= sh;
    int db = nbits[k] - is_signed[k];
    int rs = 23 + is_signed[k] - db;           /* :418-422 */
    if (!is_signed[k]) rs += 1;
    d = d + (1u << (rs - 1));                  /* :425-427 round to nearest */
    if (d > (1u << 24) - 1) d = (1u << 24) - 1; /* :429-430 */
    d >>= rs;                                  /* :434 */
    if (is_signed[k]) d |= (b & 0x80000000u) >> (31 - db); /* :435-441 */
    digits_out[k] = d & mask32(nbits[k]);
  }
}

/* reconstruct_custom_float_with_exponent, shared_exponent=true
   (codegen_llvm_quant.cpp:660-757).  The zero test uses the digits as the
   reference loads them (sign-extended for signed types), so a signed member
   whose magnitude rounded to 0 but kept its sign bit decodes to
   -2^(e - db - 1) of the group scale, exactly as the reference does. */
float orc_shared_decode(uint32_t exp, uint32_t digits, int nbits,
                        int is_signed, int exp_bits) {
  int db = nbits - is_signed;
  int32_t input_digits = is_signed ? orc_sext(digits, nbits)
                                   : (int32_t)(digits & mask32(nbits));
  uint32_t off = (uint32_t)exp_conv(exp_bits);
  if (exp_bits < 8 && exp == 0) off = 0;       /* :675-681 */
  uint32_t d = digits & mask32(nbits);         /* :688-691 */
  uint32_t sign = 0;
  if (is_signed) {                             /* :696-702 */
    sign = d & (1u << db);
    d ^= sign;
    sign <<= (31 - db);
    d <<= 1;
  }
  int lz = d ? __builtin_clz(d) : 32;          /* :706-709 ctlz, zero defined */
  int32_t extra = (31 - db) - lz;              /* :710-711 */
  off += (uint32_t)extra;                      /* :712 */
  if (!is_signed) off += 1;                    /* :714-716 */
  int32_t dshift = (23 - db) - extra;          /* :718-720 */
  d <<= dshift;
  uint32_t frac = d & 0x7FFFFFu;
  uint32_t ev = exp + off;
  uint32_t fb = (ev << 23) | frac;
  if (exp == 0 || input_digits == 0) fb = 0;   /* :738-747 */
  if (is_signed) fb |= sign;                   /* :749-754 */
  return u2f(fb);
}

/* Bit stream over u64 words, LSB first (struct_llvm.cpp:95-98: word =
   offset/64, bit = offset%64).  A field crossing a word boundary keeps its
   low bits at the top of word w and its high bits at the bottom of word w+1
   (store_split/load_split, codegen_llvm_quant.cpp:798-878). */
void orc_put_bits(uint64_t *rec, int off, int nbits, uint32_t code) {
  uint64_t m = mask64(nbits);
  uint64_t v = (uint64_t)code & m;
  int w = off >> 6, b = off & 63;
  rec[w] = (rec[w] & ~(m << b)) | (v << b);
  if (b + nbits > 64) {
    int lo = 64 - b;
    rec[w + 1] = (rec[w + 1] & ~(m >> lo)) | (v >> lo);
  }
}

uint32_t orc_get_bits(const uint64_t *rec, int off, int nbits) {
  int w = off >> 6, b = off & 63;
  uint64_t v = rec[w] >> b;
  if (b + nbits > 64) v |= rec[w + 1] << (64 - b);
  return (uint32_t)(v & mask64(nbits));
}

/* Counter-based dithering noise.  lowbias32 is the 32-bit integer hash by
   C. Wellons (hash-prospector, 2018): xorshift-multiply rounds with
   constants 0x7feb352d and 0x846ca68b. */
uint32_t orc_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

uint32_t orc_noise_key(uint64_t seed, uint32_t step, uint32_t comp) {
  uint32_t k = orc_hash32((uint32_t)seed ^ orc_hash32((uint32_t)(seed >> 32)));
  k = orc_hash32(step ^ k);
  return orc_hash32(comp ^ k);
}

float orc_noise_u(uint64_t seed, uint32_t step, uint32_t slot, uint32_t comp) {
  uint32_t h = orc_hash32(slot ^ orc_noise_key(seed, step, comp));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

/* Dithering noise: one hash per component pair (2j, 2j+1) with the key of
   j; component comp takes the low (even comp) or high (odd comp) 16 bits h16
   of it: u = (h16 + 1/2) 2^-16. */
float orc_dither_u(uint64_t seed, uint32_t step, uint32_t slot, uint32_t comp) {
  uint32_t h = orc_hash32(slot ^ orc_noise_key(seed, step, comp >> 1));
  uint32_t b = (comp & 1u) ? (h >> 16) : (h & 0xFFFFu);
  return ((float)b + 0.5f) * (1.0f / 65536.0f);
}

void orc_noise_fill(uint64_t seed, uint32_t step, int nc, int64_t n, double *xi) {
  for (int c = 0; c < nc; c++)
    for (int64_t p = 0; p < n; p++)
      xi[(int64_t)c * n + p] = (double)orc_noise_u(seed, step, (uint32_t)p, (uint32_t)c) - 0.5;
}

/* ---- record level ----------------------------------------------------- */

static int group_members(const orc_comp *c, int nc, int k, int *idx) {
  int n = 0;
  for (int j = 0; j < nc; j++)
    if (c[j].kind == ORC_FLOAT && c[j].shared &&
        c[j].exp_offset == c[k].exp_offset)
      idx[n++] = j;
  return n;
}

static void encode_group(const orc_comp *c, int nc, uint64_t *rec, int k,
                         const float *v /* indexed by component */) {
  int idx[64];
  float f[64];
  int32_t nb[64], sg[64];
  uint32_t dg[64], e;
  int n = group_members(c, nc, k, idx);
  for (int j = 0; j < n; j++) {
    f[j] = v[idx[j]];
    nb[j] = c[idx[j]].nbits;
    sg[j] = c[idx[j]].is_signed;
  }
  orc_shared_encode(f, n, nb, sg, c[k].exp_bits, &e, dg);
  orc_put_bits(rec, c[k].exp_offset, c[k].exp_bits, e);
  for (int j = 0; j < n; j++)
    orc_put_bits(rec, c[idx[j]].bit_offset, c[idx[j]].nbits, dg[j]);
}

static int first_of_group(const orc_comp *c, int k) {
  for (int j = 0; j < k; j++)
    if (c[j].kind == ORC_FLOAT && c[j].shared && c[j].exp_offset == c[k].exp_offset)
      return 0;
  return 1;
}

static void encode_one(const orc_comp *c, uint64_t *rec, float v, float u) {
  uint32_t e, d;
  switch (c->kind) {
    case ORC_FIXED:
      orc_put_bits(rec, c->bit_offset, c->nbits,
                   orc_fixed_encode_f32(v, c->scale, c->offset, c->nbits,
                                        c->is_signed, c->dither, u));
      break;
    case ORC_INT: {
      /* custom int store of a real value: fptosi/fptoui then mask */
      uint32_t q = c->is_signed ? (uint32_t)cvt_i32_f64(v) : cvt_u32_f64(v);
      orc_put_bits(rec, c->bit_offset, c->nbits, q);
      break;
    }
    case ORC_FLOAT:
      orc_float_encode(v, c->nbits, c->is_signed, c->exp_bits, &e, &d);
      orc_put_bits(rec, c->exp_offset, c->exp_bits, e);
      orc_put_bits(rec, c->bit_offset, c->nbits, d);
      break;
    default: /* ORC_F32 */
      orc_put_bits(rec, c->bit_offset, 32, f2u(v));
      break;
  }
}

static float decode_one(const orc_comp *c, const uint64_t *rec) {
  uint32_t q = orc_get_bits(rec, c->bit_offset, c->nbits);
  switch (c->kind) {
    case ORC_FIXED:
      return orc_fixed_decode_f32(q, c->nbits, c->is_signed, c->scale, c->offset);
    case ORC_INT:
      return c->is_signed ? (float)orc_sext(q, c->nbits) : (float)q;
    case ORC_FLOAT: {
      uint32_t e = orc_get_bits(rec, c->exp_offset, c->exp_bits);
      return c->shared ? orc_shared_decode(e, q, c->nbits, c->is_signed, c->exp_bits)
                       : orc_float_decode(e, q, c->nbits, c->is_signed, c->exp_bits);
    }
    default:
      return u2f(q);
  }
}

/* bit_pack_assign of a whole record (codegen_llvm_quant.cpp:884-952):
   every member is encoded and ORed into its word; words are written whole. */
void orc_pack_f32(const orc_comp *comps, int nc, int nwords, const float *vals,
                  const float *noise, uint64_t seed, uint32_t step,
                  uint64_t *words, int64_t n) {
  orc_pack_range_f32(comps, nc, nwords, vals, noise, seed, step, words, n, 0, n);
}

void orc_pack_range_f32(const orc_comp *comps, int nc, int nwords, const float *vals,
                        const float *noise, uint64_t seed, uint32_t step,
                        uint64_t *words, int64_t n, int64_t lo, int64_t hi) {
  for (int64_t p = lo; p < hi; p++) {
    uint64_t rec[64];
    float v[64];
    memset(rec, 0, sizeof(uint64_t) * (size_t)nwords);
    for (int k = 0; k < nc; k++) v[k] = vals[(int64_t)k * n + p];
    for (int k = 0; k < nc; k++) {
      const orc_comp *c = &comps[k];
      if (c->kind == ORC_FLOAT && c->shared) {
        if (first_of_group(comps, k)) encode_group(comps, nc, rec, k, v);
        continue;
      }
      float u = 0.0f;
      if (c->kind == ORC_FIXED && c->dither)
        u = noise ? noise[(int64_t)k * n + p]
                  : orc_dither_u(seed, step, (uint32_t)p, (uint32_t)k);
      encode_one(c, rec, v[k], u);
    }
    for (int w = 0; w < nwords; w++) words[(int64_t)w * n + p] = rec[w];
  }
}

void orc_unpack_f32(const orc_comp *comps, int nc, int nwords,
                    const uint64_t *words, float *vals, int64_t n) {
  orc_unpack_range_f32(comps, nc, nwords, words, vals, n, 0, n);
}

void orc_unpack_range_f32(const orc_comp *comps, int nc, int nwords,
                          const uint64_t *words, float *vals, int64_t n,
                          int64_t lo, int64_t hi) {
  for (int64_t p = lo; p < hi; p++) {
    uint64_t rec[64];
    for (int w = 0; w < nwords; w++) rec[w] = words[(int64_t)w * n + p];
    for (int k = 0; k < nc; k++) vals[(int64_t)k * n + p] = decode_one(&comps[k], rec);
  }
}

/* f64 compute: fixed-point and int members in double; custom floats and
   unquantized f32 members keep f32 (a CustomFloatType with an exponent
   computes in f32, TS/taichi/ir/type.cpp:129-155). */
static double decode_one_f64(const orc_comp *c, const uint64_t *rec) {
  uint32_t q = orc_get_bits(rec, c->bit_offset, c->nbits);
  if (c->kind == ORC_FIXED)
    return orc_fixed_decode_f64(q, c->nbits, c->is_signed, c->scale, c->offset);
  if (c->kind == ORC_INT)
    return c->is_signed ? (double)orc_sext(q, c->nbits) : (double)(q & mask32(c->nbits));
  return (double)decode_one(c, rec);
}

void orc_pack_f64(const orc_comp *comps, int nc, int nwords, const double *vals,
                  const double *noise, uint64_t seed, uint32_t step,
                  uint64_t *words, int64_t n) {
  for (int64_t p = 0; p < n; p++) {
    uint64_t rec[64];
    float v[64];
    memset(rec, 0, sizeof(uint64_t) * (size_t)nwords);
    for (int k = 0; k < nc; k++) v[k] = (float)vals[(int64_t)k * n + p];
    for (int k = 0; k < nc; k++) {
      const orc_comp *c = &comps[k];
      if (c->kind == ORC_FLOAT && c->shared) {
        if (first_of_group(comps, k)) encode_group(comps, nc, rec, k, v);
        continue;
      }
      if (c->kind == ORC_FIXED) {
        double u = 0.0;
        if (c->dither)
          u = noise ? noise[(int64_t)k * n + p]
                    : (double)orc_dither_u(seed, step, (uint32_t)p, (uint32_t)k);
        orc_put_bits(rec, c->bit_offset, c->nbits,
                     orc_fixed_encode_f64(vals[(int64_t)k * n + p], c->scale,
                                          c->offset, c->nbits, c->is_signed,
                                          c->dither, u));
      } else if (c->kind == ORC_INT) {
        const double x = vals[(int64_t)k * n + p];
        orc_put_bits(rec, c->bit_offset, c->nbits, c->is_signed ? (uint32_t)cvt_i32_f64(x) : cvt_u32_f64(x));
      } else {
        encode_one(c, rec, v[k], 0.0f);
      }
    }
    for (int w = 0; w < nwords; w++) words[(int64_t)w * n + p] = rec[w];
  }
}

void orc_unpack_f64(const orc_comp *comps, int nc, int nwords,
                    const uint64_t *words, double *vals, int64_t n) {
  for (int64_t p = 0; p < n; p++) {
    uint64_t rec[64];
    for (int w = 0; w < nwords; w++) rec[w] = words[(int64_t)w * n + p];
    for (int k = 0; k < nc; k++) vals[(int64_t)k * n + p] = decode_one_f64(&comps[k], rec);
  }
}

/* Store of a single member into an existing record.  A non-shared member is
   a masked store that leaves the other bits alone (codegen_llvm.cpp:1270-1300,
   runtime.cpp:1681-1708).  A shared-exponent member re-encodes its whole
   group, the other members being reconstructed from the record first
   (store_floats_with_shared_exponents, codegen_llvm_quant.cpp:347-383). */
void orc_store_one_f32(const orc_comp *comps, int nc, uint64_t *rec, int c,
                       float v, float u) {
  const orc_comp *cc = &comps[c];
  if (cc->kind == ORC_FLOAT && cc->shared) {
    float vals[64];
    for (int k = 0; k < nc; k++) vals[k] = decode_one(&comps[k], rec);
    vals[c] = v;
    encode_group(comps, nc, rec, c, vals);
    return;
  }
  encode_one(cc, rec, v, u);
}
