// Device-side codec for quantized particle records on gfx950.
//
// A record is held in registers as 32-bit words r[0..2*NW).  Component
// parameters are uniform (kernel arguments), so every branch on a
// component's kind/width is a scalar branch and every r[] index is an SGPR
// index.  Bit semantics follow the reference codegen (TS/ =
// taichi_submission/):
//   fixed encode   TS/taichi/codegen/codegen_llvm_quant.cpp:44-80 and
//                  TS/taichi/runtime/llvm/runtime.cpp:1732-1762
//   fixed decode   codegen_llvm_quant.cpp:624-641
//   custom float   codegen_llvm_quant.cpp:232-301 (store), 660-757 (load)
//   shared exp.    codegen_llvm_quant.cpp:347-499 (store), 660-757 (load)
//   bit placement  TS/taichi/struct/struct_llvm.cpp:71-116, split fields
//                  codegen_llvm_quant.cpp:798-878
// This file is compiled with -ffp-contract=off: each float operation of the
// encode/decode is rounded on its own, as in the reference's unfused form.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "aq_mpm.h"

namespace aq {

enum : uint8_t { K_FIXED = AQ_FIXED, K_INT = AQ_INT, K_FLOAT = AQ_FLOAT, K_F32 = AQ_F32 };

// Kernel-side component: 32 bytes, uniform (one s_load_dwordx8).
struct CompK {
  uint8_t kind, nbits, is_signed, dither;
  uint8_t exp_bits, shared, gfirst, glast;  // shared groups are contiguous runs
  uint16_t off, eoff;                        // bit offsets of digits / exponent
  uint32_t conv;                             // exponent conversion offset 127-2^(e-1)+1
  float enc_s;                               // (float)(1.0 / scale)
  float dec_s;                               // (float)scale
  float offs;                                // (float)offset
  uint32_t pad;
};
static_assert(sizeof(CompK) == 32, "CompK is one scalar load");

struct LayoutK {
  int32_t nc;
  int32_t nw;
  CompK c[AQ_MAX_COMPS];
};

// Record codec of the MPM kernels: component descriptors in device memory,
// read where they are used (see comp_at), plus per-component bit masks of
// the record's structure, which stay in scalar registers.
struct CodecK {
  const CompK *c;
  int32_t nc, nw;
  uint32_t shared_mask;  // bit c: member of a shared-exponent group
  uint32_t gfirst_mask;  // bit c: first member of its group
  uint32_t glast_mask;   // bit c: last member of its group
  uint32_t key_base;     // dithering key of component pair j: hash32(j ^ key_base)
  uint32_t stream;       // fields lie in encode order at increasing offsets (StreamSink applies)
  uint32_t pad;
  unsigned long long *wraps;  // per component: stores whose code wrapped (null: not counted)
  // f64 compute: per component {1.0 / scale, scale, offset, 0} as doubles
  // (CompK holds them rounded to float); null for f32 compute
  const double *cd;
};

// Descriptors are loaded where they are used, a few at a time: the empty asm
// hides the table pointer's value from the optimiser so the scalar loads are
// not hoisted to the kernel entry (holding 16 descriptors there overflows the
// scalar registers), and a batch of loads shares one wait.  (Through the
// constant address space the uniform loads are scalar loads.)
typedef const __attribute__((address_space(4))) uint32_t ConstU32;
constexpr int COMP_BATCH = 4;

#ifdef AQ_JIT_LAYOUT
// Run-time specialisation of a kernel to one record layout: the generated
// source defines AQ_JIT_COMPS (the CompK initialisers), AQ_JIT_SHARED,
// AQ_JIT_GFIRST, AQ_JIT_GLAST (structure masks), AQ_JIT_NW (record words)
// and AQ_JIT_G2P_MASK (words G2P reads).  Descriptors and masks are then
// compile-time constants: kinds, offsets and widths fold into the code.
namespace jit {
constexpr CompK kComps[] = AQ_JIT_COMPS;
constexpr uint32_t kShared = AQ_JIT_SHARED, kGFirst = AQ_JIT_GFIRST, kGLast = AQ_JIT_GLAST;
constexpr int kNW = AQ_JIT_NW;
constexpr uint32_t kG2PMask = AQ_JIT_G2P_MASK;
}  // namespace jit

__device__ __forceinline__ ConstU32 *comp_table(const CodecK &) { return nullptr; }
__device__ __forceinline__ CompK comp_load(ConstU32 *, int c) { return jit::kComps[c]; }
#else
__device__ __forceinline__ ConstU32 *comp_table(const CodecK &cod) {
  ConstU32 *p = (ConstU32 *)cod.c;
  asm volatile("" : "+s"(p));
  return p;
}

__device__ __forceinline__ CompK comp_load(ConstU32 *p, int c) {
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = p[c * 8 + i];
  CompK k;
  __builtin_memcpy(&k, w, sizeof(k));
  return k;
}
#endif

__device__ __forceinline__ CompK comp_at(const CodecK &cod, int c) { return comp_load(comp_table(cod), c); }

__device__ __forceinline__ uint32_t lowmask(uint32_t nb) {
  return nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);
}

__device__ __forceinline__ int32_t sext(uint32_t q, uint32_t nb) {
  const uint32_t sh = 32u - nb;
  return (int32_t)(q << sh) >> sh;
}

// Field [off, off+nb) of the record, nb <= 32; a field may cross one 32-bit
// word boundary (and so one u64 word boundary).
template <int NW32>
__device__ __forceinline__ uint32_t get_field(const uint32_t (&r)[NW32], uint32_t off,
                                              uint32_t nb) {
  const uint32_t i = off >> 5, b = off & 31u;
  uint32_t v;
  if (b + nb > 32u) {
    v = __builtin_amdgcn_alignbit(r[i + 1], r[i], b);
  } else {
    v = r[i] >> b;
  }
  return v & lowmask(nb);
}

// Records are encoded in a per-thread row of LDS words (rows of an odd
// number of words, so the 64 lanes of a wavefront hit distinct banks): a
// field at a uniform bit offset becomes one or two ds_or_b32 at a uniform
// word index, instead of a read-modify-write of a register array at a
// runtime index.
// OR a field into a row whose field bits are zero.
__device__ __forceinline__ void row_put(uint32_t *row, uint32_t off, uint32_t nb, uint32_t v) {
  const uint32_t i = off >> 5, b = off & 31u;
  v &= lowmask(nb);
  atomicOr(row + i, v << b);
  if (b + nb > 32u) atomicOr(row + i + 1, v >> (32u - b));
}

// Replace a field of a row (masked store).
__device__ __forceinline__ void row_set(uint32_t *row, uint32_t off, uint32_t nb, uint32_t v) {
  const uint32_t i = off >> 5, b = off & 31u;
  const uint32_t m = lowmask(nb);
  v &= m;
  row[i] = (row[i] & ~(m << b)) | (v << b);
  if (b + nb > 32u) {
    const uint32_t s = 32u - b;
    row[i + 1] = (row[i + 1] & ~(m >> s)) | (v >> s);
  }
}

// fptosi / fptoui with saturation and NaN -> 0 (the reference's values are
// LLVM poison outside the int range): the hardware conversions
// v_cvt_i32_f32 / v_cvt_u32_f32 truncate toward zero, clamp to the integer
// range and turn NaN into 0 (CDNA3 ISA, VOP1 conversions), which is exactly
// this semantics, in one instruction.  (The C conversion would let the
// compiler treat out-of-range inputs as undefined.)
__device__ __forceinline__ uint32_t cvt_sat(float y, bool is_signed) {
  uint32_t r;
  if (is_signed)
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(y));
  else
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(y));
  return r;
}

// lowbias32 (C. Wellons, hash-prospector): xorshift-multiply rounds with
// constants 0x7feb352d and 0x846ca68b.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// Uniform noise in [0, 1) on a 2^-24 grid: the hash of the record slot xor a
// per-(seed, step, component) key (the modelled rounding noise of the
// adjoint run).
__device__ __forceinline__ float noise_u(uint32_t slot, uint32_t key) {
  return (float)(hash32(slot ^ key) >> 8) * (1.0f / 16777216.0f);
}
// Dithering noise of component c: one hash serves the pair of components
// (2j, 2j+1), key_j = hash32(j ^ key_base); component c takes half (c & 1)
// of it, u = (h16 + 1/2) 2^-16, exact, in (0, 1) with mean 1/2.
__device__ __forceinline__ float dither_u(uint32_t slot, uint32_t pair_key, int c) {
  const uint32_t h = hash32(slot ^ pair_key);
  const uint32_t b = (c & 1) ? (h >> 16) : (h & 0xFFFFu);
  return __uint_as_float(0x3F800040u | (b << 7)) - 1.0f;
}

// Fixed-point code of v before masking to its width: the saturated integer
// (cvt_sat) of y, the rounded scaled value it truncates.
__device__ __forceinline__ uint32_t enc_fixed_sat(float v, const CompK &c, float u, float &y) {
  y = v - c.offs;
  y = y * c.enc_s;
  if (c.dither) {
    y = y + u;
    y = y - 0.5f;
  }
  y = y + __builtin_copysignf(0.5f, y);
  return cvt_sat(y, c.is_signed);
}

__device__ __forceinline__ uint32_t enc_fixed(float v, const CompK &c, float u) {
  float y;
  return enc_fixed_sat(v, c, u, y) & lowmask(c.nbits);
}

// Whether the code of a fixed or int member wraps: trunc(y) outside the
// nb-bit range (the stored code then aliases another value), or y is NaN.
// q is cvt_sat(y).
__device__ __forceinline__ bool code_wraps(float y, uint32_t q, uint32_t nb, bool is_signed) {
  if (y != y) return true;
  if (is_signed) {
    if (nb >= 32u) return y >= 2147483648.0f || y < -2147483648.0f;
    const uint32_t sh = 32u - nb;
    return (int32_t)(q << sh) >> sh != (int32_t)q;
  }
  if (y <= -1.0f) return true;
  if (nb >= 32u) return y >= 4294967296.0f;
  return (q >> nb) != 0u;
}

__device__ __forceinline__ float dec_fixed(uint32_t q, const CompK &c) {
  const float cf = c.is_signed ? (float)sext(q, c.nbits) : (float)q;
  const float t = cf * c.dec_s;
  // cf is never -0 and dec_s > 0 (checked at layout time), so t is never
  // -0 and t + 0 == t: a zero offset needs no add
  return c.offs == 0.0f ? t : t + c.offs;
}

// ---- f64 compute: fixed-point and int members encoded and decoded in
// double (float_to_custom_int / reconstruct_custom_float with compute type
// f64, codegen_llvm_quant.cpp:44-80, 624-641); custom floats keep f32
// compute (a CustomFloatType with an exponent computes in f32,
// TS/taichi/ir/type.cpp:129-155).
// v_cvt_i32_f64 / v_cvt_u32_f64 truncate toward zero, clamp to the integer
// range and turn NaN into 0, as cvt_sat does for f32.
__device__ __forceinline__ uint32_t cvt_sat_f64(double y, bool is_signed) {
  uint32_t r;
  if (is_signed)
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(y));
  else
    asm("v_cvt_u32_f64 %0, %1" : "=v"(r) : "v"(y));
  return r;
}

// y = (v - offset) * (1 / scale) [+ u - 1/2] + copysign(1/2, .), all in
// double; returns the saturated integer of y.
__device__ __forceinline__ uint32_t enc_fixed_sat_f64(double v, const CompK &c, const double *cd, double u,
                                                      double &y) {
  y = v - cd[2];
  y = y * cd[0];
  if (c.dither) {
    y = y + u;
    y = y - 0.5;
  }
  y = y + __builtin_copysign(0.5, y);
  return cvt_sat_f64(y, c.is_signed);
}

__device__ __forceinline__ double dec_fixed_f64(uint32_t q, const CompK &c, const double *cd) {
  const double cf = c.is_signed ? (double)sext(q, c.nbits) : (double)(q & lowmask(c.nbits));
  const double t = cf * cd[1];
  return cd[2] == 0.0 ? t : t + cd[2];
}

__device__ __forceinline__ bool code_wraps_f64(double y, uint32_t q, uint32_t nb, bool is_signed) {
  if (y != y) return true;
  if (is_signed) {
    if (nb >= 32u) return y >= 2147483648.0 || y < -2147483648.0;
    const uint32_t sh = 32u - nb;
    return (int32_t)(q << sh) >> sh != (int32_t)q;
  }
  if (y <= -1.0) return true;
  if (nb >= 32u) return y >= 4294967296.0;
  return (q >> nb) != 0u;
}

__device__ __forceinline__ uint32_t f32_exp(float f) {
  return (__float_as_uint(f) >> 23) & 0xFFu;
}

// Custom float with its own exponent: returns digits, writes the exponent.
__device__ __forceinline__ uint32_t enc_float(float f, const CompK &c, uint32_t &e_out) {
  const uint32_t db = c.nbits - c.is_signed;
  uint32_t b = __float_as_uint(f);
  if (db < 23u) b += 1u << (22u - db);
  const uint32_t e = (b >> 23) & 0xFFu;
  uint32_t d = (uint32_t)((int32_t)b >> (23u - db)) & lowmask(db);
  if (c.is_signed) d |= (b & 0x80000000u) >> (31u - db);
  int32_t es = (int32_t)e - (e ? (int32_t)c.conv : 0);
  es = es < 0 ? 0 : es;
  e_out = (uint32_t)es & lowmask(c.exp_bits);
  return es ? (d & lowmask(c.nbits)) : 0u;
}

__device__ __forceinline__ float dec_float(uint32_t e, uint32_t q, const CompK &c) {
  const uint32_t db = c.nbits - c.is_signed;
  const uint32_t off = (c.exp_bits < 8 && e == 0u) ? 0u : c.conv;
  const uint32_t d = (q & lowmask(c.nbits)) << (23u - db);
  uint32_t fb = ((e + off) << 23) | (d & 0x7FFFFFu);
  if (c.is_signed) fb |= (d & 0x800000u) << 8;
  return __uint_as_float(fb);
}

// Digits of one member of a shared-exponent group with max f32 exponent m.
__device__ __forceinline__ uint32_t enc_shared(float f, uint32_t m, const CompK &c) {
  const uint32_t b = __float_as_uint(f);
  const uint32_t e = (b >> 23) & 0xFFu;
  uint32_t sh = m - e;
  sh = sh > 31u ? 31u : sh;
  uint32_t d = ((b & 0x7FFFFFu) | 0x800000u) >> sh;
  const uint32_t db = c.nbits - c.is_signed;
  const uint32_t rs = 24u - db;  // 23 + signed - db, +1 when unsigned (:418-422)
  d = d + (1u << (rs - 1u));
  d = d > 0xFFFFFFu ? 0xFFFFFFu : d;
  d >>= rs;
  if (c.is_signed) d |= (b & 0x80000000u) >> (31u - db);
  return d & lowmask(c.nbits);
}

__device__ __forceinline__ uint32_t shared_exp_field(uint32_t m, const CompK &c) {
  int32_t es = (int32_t)m - (m ? (int32_t)c.conv : 0);
  es = es < 0 ? 0 : es;
  return (uint32_t)es & lowmask(c.exp_bits);
}

__device__ __forceinline__ float dec_shared(uint32_t e, uint32_t q, const CompK &c) {
  const uint32_t db = c.nbits - c.is_signed;
  q &= lowmask(c.nbits);
  // zero test on the digits as loaded (sign-extended for signed members)
  const bool zero_digits = c.is_signed ? (sext(q, c.nbits) == 0) : (q == 0u);
  uint32_t off = (c.exp_bits < 8 && e == 0u) ? 0u : c.conv;
  uint32_t d = q, sign = 0u;
  if (c.is_signed) {
    sign = d & (1u << db);
    d ^= sign;
    sign <<= (31u - db);
    d <<= 1;
  }
  const int32_t lz = d ? __builtin_clz(d) : 32;
  const int32_t extra = (int32_t)(31u - db) - lz;
  off += (uint32_t)extra;
  if (!c.is_signed) off += 1u;
  d <<= (uint32_t)((int32_t)(23u - db) - extra);
  uint32_t fb = ((e + off) << 23) | (d & 0x7FFFFFu);
  if (e == 0u || zero_digits) fb = 0u;
  return __uint_as_float(fb | sign);
}

// dec_shared of the encoder's own output, for signed members: when the
// group's unit digit 2^(e + conv - 126 - db) and every value it scales are
// normal floats (e + conv + 1 - db >= 3 and e + conv <= 254), the decode is
// the digit magnitude times that unit (a sign-only pattern decodes to a
// quarter unit, as dec_shared's renormalisation gives); otherwise
// dec_shared.  About half the instructions of dec_shared per member.
__device__ __forceinline__ float dec_shared_enc(uint32_t e, uint32_t q, const CompK &c) {
  if (c.is_signed) {
    const uint32_t db = c.nbits - 1u;
    const int32_t ue = (int32_t)(e + c.conv) + 1 - (int32_t)db;
    if (e != 0u && ue >= 3 && e + c.conv <= 254u) {
      q &= lowmask(c.nbits);
      const uint32_t mag = q & lowmask(db), sign = (q >> db) & 1u;
      const float m = mag ? (float)mag : (sign ? 0.25f : 0.0f);
      const float r = m * __uint_as_float((uint32_t)ue << 23);
      return __uint_as_float(__float_as_uint(r) | (sign << 31));
    }
  }
  return dec_shared(e, q, c);
}

// Decode a member given a field reader get(off, nbits).
template <class G>
__device__ __forceinline__ float decode_with(G &&get, const CompK &c) {
  if (c.kind == K_FIXED) return dec_fixed(get(c.off, c.nbits), c);
  if (c.kind == K_F32) return __uint_as_float(get(c.off, 32u));
  if (c.kind == K_INT) {
    const uint32_t q = get(c.off, c.nbits);
    return c.is_signed ? (float)sext(q, c.nbits) : (float)q;
  }
  const uint32_t e = get(c.eoff, c.exp_bits);
  const uint32_t q = get(c.off, c.nbits);
  return c.shared ? dec_shared(e, q, c) : dec_float(e, q, c);
}

// The same in f64 compute: cd = CodecK::cd + 4 c.
template <class G>
__device__ __forceinline__ double decode_with_f64(G &&get, const CompK &c, const double *cd) {
  if (c.kind == K_FIXED) return dec_fixed_f64(get(c.off, c.nbits), c, cd);
  if (c.kind == K_INT) {
    const uint32_t q = get(c.off, c.nbits);
    return c.is_signed ? (double)sext(q, c.nbits) : (double)q;
  }
  return (double)decode_with(get, c);
}

// Decode component c of a record held in registers.
template <int NW32>
__device__ __forceinline__ float decode_comp(const uint32_t (&r)[NW32], const CompK &c) {
  return decode_with([&](uint32_t off, uint32_t nb) { return get_field(r, off, nb); }, c);
}

// Field of a record held in an LDS row (one or two ds_read_b32 at a uniform index).
__device__ __forceinline__ uint32_t row_get(const uint32_t *row, uint32_t off, uint32_t nb) {
  const uint32_t i = off >> 5, b = off & 31u;
  const uint32_t v = (b + nb > 32u) ? __builtin_amdgcn_alignbit(row[i + 1], row[i], b) : (row[i] >> b);
  return v & lowmask(nb);
}

__device__ __forceinline__ float decode_row(const uint32_t *row, const CompK &c) {
  return decode_with([&](uint32_t off, uint32_t nb) { return row_get(row, off, nb); }, c);
}

__device__ __forceinline__ double decode_row_f64(const uint32_t *row, const CompK &c, const double *cd) {
  return decode_with_f64([&](uint32_t off, uint32_t nb) { return row_get(row, off, nb); }, c, cd);
}

// Encode one member (fixed, own-exponent float, f32 or int) into a row.
__device__ __forceinline__ void encode_member(float v, const CompK &k, float u, uint32_t *row, bool masked) {
  if (k.kind == K_FIXED) {
    const uint32_t q = enc_fixed(v, k, u);
    if (masked) row_set(row, k.off, k.nbits, q); else row_put(row, k.off, k.nbits, q);
  } else if (k.kind == K_FLOAT) {
    uint32_t e;
    const uint32_t d = enc_float(v, k, e);
    if (masked) {
      row_set(row, k.eoff, k.exp_bits, e);
      row_set(row, k.off, k.nbits, d);
    } else {
      row_put(row, k.eoff, k.exp_bits, e);
      row_put(row, k.off, k.nbits, d);
    }
  } else if (k.kind == K_F32) {
    if (masked) row_set(row, k.off, 32u, __float_as_uint(v)); else row_put(row, k.off, 32u, __float_as_uint(v));
  } else {
    const uint32_t q = cvt_sat(v, k.is_signed);
    if (masked) row_set(row, k.off, k.nbits, q); else row_put(row, k.off, k.nbits, q);
  }
}

// Field writers of the record encoder: put(off, nbits, value) with value
// masked to nbits.
// RowSink ORs fields into a zeroed per-thread LDS row (any layout).
struct RowSink {
  uint32_t *row;
  __device__ __forceinline__ void put(uint32_t off, uint32_t nb, uint32_t v) { row_put(row, off, nb, v); }
};

// StreamSink builds the record in a 64-bit register and stores each word of
// the plane-major record as it completes (words[i * cap + p]); fields must
// come at non-decreasing offsets (CodecK::stream).  Every offset is uniform,
// so the word bookkeeping is scalar and no LDS is touched.
struct StreamSink {
  uint64_t *words;
  int64_t cap, p;
  uint32_t nw;
  uint32_t pos;  // bit position of the stream (uniform)
  uint64_t acc;  // bits of the current word
  __device__ __forceinline__ void store() {
    __builtin_nontemporal_store(acc, words + (int64_t)(pos >> 6) * cap + p);
  }
  __device__ __forceinline__ void put(uint32_t off, uint32_t nb, uint32_t v) {
    while ((pos >> 6) < (off >> 6)) {  // words before the field are complete
      store();
      acc = 0ull;
      pos = ((pos >> 6) + 1u) << 6;
    }
    pos = off;
    const uint32_t b = pos & 63u;
    acc |= (uint64_t)v << b;
    if (b + nb >= 64u) {
      store();
      acc = b + nb > 64u ? (uint64_t)v >> (64u - b) : 0ull;
    }
    pos += nb;
  }
  __device__ __forceinline__ void finish() {
    while ((pos >> 6) < nw) {
      store();
      acc = 0ull;
      pos = ((pos >> 6) + 1u) << 6;
    }
  }
};

// RegSink ORs fields into NWD 64-bit registers; with compile-time offsets
// (a specialised layout) every index is a constant.
template <int NWD> struct RegSink {
  uint64_t w[NWD];
  __device__ __forceinline__ void put(uint32_t off, uint32_t nb, uint32_t v) {
    const uint32_t i = off >> 6, b = off & 63u;
    w[i] |= (uint64_t)v << b;
    if (b + nb > 64u) w[i + 1] |= (uint64_t)v >> (64u - b);
  }
};

// Encode NC values (compile-time count) with `sink` and return in dec[] the
// values the stored fields decode to (the decode of each field's digits,
// computed from the digits in registers).  Fields are written in component
// order, a shared group's exponent before its first member and an own
// exponent before its digits.
template <int NC, class Sink>
__device__ __forceinline__ void encode_record(const float (&v)[NC], const CodecK &cod, uint32_t slot,
                                              Sink &sink, float (&dec)[NC]) {
#ifdef AQ_JIT_LAYOUT
  constexpr uint32_t sm = jit::kShared, gf = jit::kGFirst, gl = jit::kGLast;
  const uint32_t kb = cod.key_base;
#else
  // the structure masks and the key base are re-read here (hidden from the
  // optimiser) rather than kept live across the caller's loops
  uint32_t sm = cod.shared_mask, gf = cod.gfirst_mask, gl = cod.glast_mask, kb = cod.key_base;
  asm volatile("" : "+s"(sm), "+s"(gf), "+s"(gl), "+s"(kb));
#endif
  // largest f32 exponent of each shared group, spread over its members
  uint32_t gmax[NC];
  uint32_t run = 0u;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    gmax[c] = 0u;
    if ((sm >> c) & 1u) {
      const uint32_t e = f32_exp(v[c]);
      run = ((gf >> c) & 1u) ? e : (run > e ? run : e);
      gmax[c] = run;
    }
  }
#pragma unroll
  for (int c = NC - 1; c >= 0; --c) {
    if ((sm >> c) & 1u) {
      run = ((gl >> c) & 1u) ? gmax[c] : run;
      gmax[c] = run;
    }
  }
  const bool count_wraps = cod.wraps != nullptr;  // uniform
  uint32_t wrapped = 0u;
  ConstU32 *table = nullptr;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c % COMP_BATCH == 0) table = comp_table(cod);
    const CompK k = comp_load(table, c);
    if ((sm >> c) & 1u) {
      const uint32_t e = shared_exp_field(gmax[c], k);
      if ((gf >> c) & 1u) sink.put(k.eoff, k.exp_bits, e);
      const uint32_t d = enc_shared(v[c], gmax[c], k);
      sink.put(k.off, k.nbits, d);
      dec[c] = dec_shared_enc(e, d, k);
    } else if (k.kind == K_FIXED) {
      // both members of a component pair hash the same (slot, key): one hash per pair
      const float u = k.dither ? dither_u(slot, hash32((uint32_t)c >> 1 ^ kb), c) : 0.0f;
      float y;
      const uint32_t qs = enc_fixed_sat(v[c], k, u, y);
      const uint32_t q = qs & lowmask(k.nbits);
      if (count_wraps && code_wraps(y, qs, k.nbits, k.is_signed)) wrapped |= 1u << c;
      sink.put(k.off, k.nbits, q);
      dec[c] = dec_fixed(q, k);
    } else if (k.kind == K_FLOAT) {
      uint32_t e;
      const uint32_t d = enc_float(v[c], k, e);
      sink.put(k.eoff, k.exp_bits, e);
      sink.put(k.off, k.nbits, d);
      dec[c] = dec_float(e, d, k);
    } else if (k.kind == K_F32) {
      sink.put(k.off, 32u, __float_as_uint(v[c]));
      dec[c] = v[c];
    } else {
      const uint32_t qs = cvt_sat(v[c], k.is_signed);
      const uint32_t q = qs & lowmask(k.nbits);
      if (count_wraps && code_wraps(v[c], qs, k.nbits, k.is_signed)) wrapped |= 1u << c;
      sink.put(k.off, k.nbits, q);
      dec[c] = k.is_signed ? (float)sext(q, k.nbits) : (float)q;
    }
  }
  // range-violation counter: the members whose code wrapped (rare; the
  // whole test is skipped when no counter is set)
  if (count_wraps)
    for (uint32_t m = wrapped; m; m &= m - 1u) atomicAdd(cod.wraps + __builtin_ctz(m), 1ull);
}

#ifndef AQ_JIT_LAYOUT
// encode_record in f64 compute: fixed-point and int members from the double
// values (dithering u as in f32 compute, exact in double), unquantized f32
// members and custom floats from the values rounded to float (custom floats
// compute in f32).  dec[] gets the decoded values in double.
template <int NC, class Sink>
__device__ __forceinline__ void encode_record_f64(const double (&v)[NC], const CodecK &cod, uint32_t slot,
                                                  Sink &sink, double (&dec)[NC]) {
  uint32_t sm = cod.shared_mask, gf = cod.gfirst_mask, gl = cod.glast_mask, kb = cod.key_base;
  asm volatile("" : "+s"(sm), "+s"(gf), "+s"(gl), "+s"(kb));
  uint32_t gmax[NC];
  uint32_t run = 0u;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    gmax[c] = 0u;
    if ((sm >> c) & 1u) {
      const uint32_t e = f32_exp((float)v[c]);
      run = ((gf >> c) & 1u) ? e : (run > e ? run : e);
      gmax[c] = run;
    }
  }
#pragma unroll
  for (int c = NC - 1; c >= 0; --c) {
    if ((sm >> c) & 1u) {
      run = ((gl >> c) & 1u) ? gmax[c] : run;
      gmax[c] = run;
    }
  }
  const bool count_wraps = cod.wraps != nullptr;
  uint32_t wrapped = 0u;
  ConstU32 *table = nullptr;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c % COMP_BATCH == 0) table = comp_table(cod);
    const CompK k = comp_load(table, c);
    const double *cd = cod.cd + 4 * c;
    if ((sm >> c) & 1u) {
      const uint32_t e = shared_exp_field(gmax[c], k);
      if ((gf >> c) & 1u) sink.put(k.eoff, k.exp_bits, e);
      const uint32_t d = enc_shared((float)v[c], gmax[c], k);
      sink.put(k.off, k.nbits, d);
      dec[c] = (double)dec_shared_enc(e, d, k);
    } else if (k.kind == K_FIXED) {
      const double u = k.dither ? (double)dither_u(slot, hash32((uint32_t)c >> 1 ^ kb), c) : 0.0;
      double y;
      const uint32_t qs = enc_fixed_sat_f64(v[c], k, cd, u, y);
      const uint32_t q = qs & lowmask(k.nbits);
      if (count_wraps && code_wraps_f64(y, qs, k.nbits, k.is_signed)) wrapped |= 1u << c;
      sink.put(k.off, k.nbits, q);
      dec[c] = dec_fixed_f64(q, k, cd);
    } else if (k.kind == K_FLOAT) {
      uint32_t e;
      const uint32_t d = enc_float((float)v[c], k, e);
      sink.put(k.eoff, k.exp_bits, e);
      sink.put(k.off, k.nbits, d);
      dec[c] = (double)dec_float(e, d, k);
    } else if (k.kind == K_F32) {
      const float f = (float)v[c];
      sink.put(k.off, 32u, __float_as_uint(f));
      dec[c] = (double)f;
    } else {
      const uint32_t qs = cvt_sat_f64(v[c], k.is_signed);
      const uint32_t q = qs & lowmask(k.nbits);
      if (count_wraps && code_wraps_f64(v[c], qs, k.nbits, k.is_signed)) wrapped |= 1u << c;
      sink.put(k.off, k.nbits, q);
      dec[c] = k.is_signed ? (double)sext(q, k.nbits) : (double)q;
    }
  }
  if (count_wraps)
    for (uint32_t m = wrapped; m; m &= m - 1u) atomicAdd(cod.wraps + __builtin_ctz(m), 1ull);
}
#endif

}  // namespace aq
