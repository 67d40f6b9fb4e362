/*
 * CPU oracle for the quantized MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference algorithm of
 * Hanke98/AutoQuantizer (a patched Taichi v0.8.4) for the quantized MPM
 * substep: the fixed-point / custom-float / shared-exponent codec, the
 * bit_pack stream, and the P2G / grid update / G2P physics.  It is the
 * checker for the HIP implementation in autoquantizer_amd/: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it follows.  Paths are
 * relative to the reference root; TS/ = taichi_submission/.
 *
 * Choices where the reference is not bit-defined (all documented in the
 * design notes of the repository):
 *   - arithmetic is unfused (the reference's fast_math=False semantics,
 *     TS/taichi/backends/cpu/jit_cpu.cpp:205-213); build with
 *     -ffp-contract=off;
 *   - float->int conversion outside the int32/uint32 range (LLVM poison in
 *     the reference) saturates, NaN converts to 0;
 *   - a custom-float exponent above 2^e-1 is masked to its field instead of
 *     spilling into neighbouring bits (TS/taichi/codegen/codegen_llvm_quant.cpp:282-284);
 *   - bit_pack_assign never writes the word past the record when the packed
 *     size is a multiple of 64 (codegen_llvm_quant.cpp:920-950);
 *   - dithering noise comes from an explicit array or from orc_dither_u(), a
 *     counter-based hash, instead of the per-thread xorshift state
 *     (TS/taichi/runtime/llvm/runtime.cpp:1560-1589).
 *
 * Parity pinning: see tests/test_oracle_*.py.  Each function below is
 * checked against the reference's known-answer tests, brute force over all
 * codes, or closed forms.
 */
#ifndef AQ_ORACLE_H
#define AQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FIXED = 0, ORC_INT = 1, ORC_FLOAT = 2, ORC_F32 = 3 };

/* One scalar component of a particle record. */
typedef struct orc_comp {
  int32_t kind;        /* ORC_FIXED / ORC_INT / ORC_FLOAT / ORC_F32          */
  int32_t nbits;       /* digit bits including the sign bit (32 for ORC_F32) */
  int32_t is_signed;
  int32_t bit_offset;  /* absolute bit offset in the record (LSB first)      */
  int32_t exp_bits;    /* ORC_FLOAT: exponent field width                     */
  int32_t exp_offset;  /* ORC_FLOAT: absolute bit offset of the exponent      */
  int32_t shared;      /* ORC_FLOAT: exponent shared by all members with the
                          same exp_offset                                     */
  int32_t dither;      /* ORC_FIXED: dithered rounding                         */
  double scale;        /* ORC_FIXED: range/2^(nbits-1) signed, range/2^nbits  */
  double offset;       /* ORC_FIXED: offset                                    */
} orc_comp;

/* ---- scalar codec ---------------------------------------------------- */
uint32_t orc_fixed_encode_f32(float v, double scale, double offset, int nbits,
                              int is_signed, int dither, float u);
uint32_t orc_fixed_encode_f64(double v, double scale, double offset, int nbits,
                              int is_signed, int dither, double u);
float orc_fixed_decode_f32(uint32_t code, int nbits, int is_signed,
                           double scale, double offset);
double orc_fixed_decode_f64(uint32_t code, int nbits, int is_signed,
                            double scale, double offset);
void orc_float_encode(float f, int nbits, int is_signed, int exp_bits,
                      uint32_t *exp_out, uint32_t *digits_out);
float orc_float_decode(uint32_t exp, uint32_t digits, int nbits, int is_signed,
                       int exp_bits);
void orc_shared_encode(const float *f, int n, const int32_t *nbits,
                       const int32_t *is_signed, int exp_bits,
                       uint32_t *exp_out, uint32_t *digits_out);
float orc_shared_decode(uint32_t exp, uint32_t digits, int nbits,
                        int is_signed, int exp_bits);

/* ---- bit stream ------------------------------------------------------- */
void orc_put_bits(uint64_t *rec, int off, int nbits, uint32_t code);
uint32_t orc_get_bits(const uint64_t *rec, int off, int nbits);
int32_t orc_sext(uint32_t code, int nbits);

/* ---- noise: counter-based hash ----------------------------------------- */
uint32_t orc_hash32(uint32_t x);
uint32_t orc_noise_key(uint64_t seed, uint32_t step, uint32_t comp);
/* uniform on a 2^-24 grid in [0, 1): hash of slot ^ key(comp) */
float orc_noise_u(uint64_t seed, uint32_t step, uint32_t slot, uint32_t comp);
/* dithering noise: (h16 + 1/2) 2^-16 from half (comp & 1) of the hash of
   slot ^ key(comp >> 1), one hash per component pair */
float orc_dither_u(uint64_t seed, uint32_t step, uint32_t slot, uint32_t comp);
/* xi[c][p] = orc_noise_u(seed, step, p, c) - 1/2, the modelled rounding
   noise of the adjoint error-gradient run (uniform in [-1/2, 1/2)). */
void orc_noise_fill(uint64_t seed, uint32_t step, int nc, int64_t n, double *xi);

/* ---- record level: words are word planes, words[w * n + p] ----------- */
/* noise: [nc][n] uniform [0,1) values, or NULL to use orc_dither_u(seed, step, p, c). */
void orc_pack_f32(const orc_comp *comps, int nc, int nwords, const float *vals,
                  const float *noise, uint64_t seed, uint32_t step,
                  uint64_t *words, int64_t n);
void orc_unpack_f32(const orc_comp *comps, int nc, int nwords,
                    const uint64_t *words, float *vals, int64_t n);
void orc_pack_f64(const orc_comp *comps, int nc, int nwords, const double *vals,
                  const double *noise, uint64_t seed, uint32_t step,
                  uint64_t *words, int64_t n);
void orc_unpack_f64(const orc_comp *comps, int nc, int nwords,
                    const uint64_t *words, double *vals, int64_t n);
void orc_pack_range_f32(const orc_comp *comps, int nc, int nwords, const float *vals,
                        const float *noise, uint64_t seed, uint32_t step,
                        uint64_t *words, int64_t n, int64_t lo, int64_t hi);
void orc_unpack_range_f32(const orc_comp *comps, int nc, int nwords,
                          const uint64_t *words, float *vals, int64_t n,
                          int64_t lo, int64_t hi);
/* single-component store into an existing record (masked, like
   TS/taichi/codegen/codegen_llvm.cpp:1270-1300 + codegen_llvm_quant.cpp:347-457) */
void orc_store_one_f32(const orc_comp *comps, int nc, uint64_t *rec, int c,
                       float v, float u);

#ifdef __cplusplus
}
#endif
#endif
