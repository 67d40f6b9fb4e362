"""Pins the oracle codec against the reference's own known-answer tests.

Each test transcribes one reference test from
taichi_submission/tests/python/ (file:line cited) into calls on the C
oracle.  Field placement follows the reference's layout rules
(TS/taichi/struct/struct_llvm.cpp:71-116: bit_pack offsets accumulate in
placement order, word = offset/64; bit_struct members share one word) and
shared exponents are placed before their first user
(TS/taichi/program/snode_expr_utils.cpp:44-66).
"""
import math

import numpy as np
import pytest
from pytest import approx


def fixed(nb, rng=1.0, signed=True, off=0, offset=0.0, dither=False):
    from oracle import orc
    return dict(kind=orc.FIXED, nbits=nb, is_signed=int(signed), bit_offset=off,
                scale=orc.fixed_scale(nb, rng, signed), offset=offset, dither=int(dither))


def cint(nb, signed=False, off=0):
    from oracle import orc
    return dict(kind=orc.INT, nbits=nb, is_signed=int(signed), bit_offset=off, scale=1.0)


def cfloat(nb, signed, exp_bits, off, exp_off, shared=False):
    from oracle import orc
    return dict(kind=orc.FLOAT, nbits=nb, is_signed=int(signed), bit_offset=off,
                exp_bits=exp_bits, exp_offset=exp_off, shared=int(shared), scale=1.0)


def roundtrip(orc, comps, nwords, values, f64=False):
    vals = np.asarray(values, dtype=np.float64 if f64 else np.float32).reshape(len(comps), -1)
    w = orc.pack(comps, nwords, vals, f64=f64)
    return orc.unpack(comps, nwords, w, f64=f64), w


# --- test_bit_pack_assign.py ------------------------------------------------

def test_bit_pack_assign(orc):
    """test_bit_pack_assign.py:10-37: fixed21 range 2 x3 + int15 in one bit_pack."""
    comps = [fixed(21, 2.0, off=0), fixed(21, 2.0, off=21), fixed(21, 2.0, off=42),
             cint(15, off=63)]
    vals = np.array([[1.5, -1.1], [1.2, 1.3], [1.3, -1.2], [15, 16]], dtype=np.float32)
    out, words = roundtrip(orc, comps, 2, vals)
    assert out[:3] == approx(vals[:3], rel=1e-4)
    assert list(out[3]) == [15, 16]
    # y straddles the word boundary: 1 bit in word 0, 14 bits in word 1
    assert (words[0, 0] >> 63) == (15 & 1)
    assert words[1, 0] == 15 >> 1


def test_bit_pack_assign_split_exactly_on_64(orc):
    """test_bit_pack_assign.py:40-69: 4 x fixed16 fill word 0 exactly, int15 starts word 1."""
    comps = [fixed(16, 2.0, off=16 * i) for i in range(4)] + [cint(15, off=64)]
    vals = np.array([[1.5, -1.1], [1.2, 1.3], [1.3, -1.2], [1.6, -1.6], [15, 16]], dtype=np.float32)
    out, words = roundtrip(orc, comps, 2, vals)
    assert out[:4] == approx(vals[:4], rel=1e-4)
    assert list(out[4]) == [15, 16]
    # the record is exactly two words: nothing is written past it
    assert words.shape == (2, 2)


# --- test_bit_pack.py / test_bit_struct.py -----------------------------------

@pytest.mark.parametrize("container", ["bit_pack", "bit_struct32"])
def test_simple_array(orc, container):
    """test_bit_pack.py:7-33 and test_bit_struct.py:7-33: ci13 = -2^i, cu19 = 2^i - 1."""
    comps = [cint(13, True, 0), cint(19, False, 13)]
    n = 12
    vals = np.array([[-(2 ** i) for i in range(n)], [2 ** i - 1 for i in range(n)]], dtype=np.float32)
    out, _ = roundtrip(orc, comps, 1, vals)
    assert np.array_equal(out, vals)


def test_bit_pack_int_load_store(orc):
    """test_bit_pack.py:38-77: ci31, cu29, ci8 in one bit_pack (68 bits, split)."""
    comps = [cint(31, True, 0), cint(29, False, 31), cint(8, True, 60)]
    cases = np.array([[2 ** 30 - 1, 2 ** 29 - 1, -(2 ** 7)], [2 ** 30 - 1, 2 ** 29 - 1, -(2 ** 7 - 1)],
                      [0, 0, 0], [123, 4567, 8], [10, 31, 11]], dtype=np.int64).T
    w = np.zeros((2, cases.shape[1]), dtype=np.uint64)
    for p in range(cases.shape[1]):
        rec = np.zeros(2, dtype=np.uint64)
        for c, comp in enumerate(comps):
            orc.lib().orc_put_bits(orc._ptr(rec), comp["bit_offset"], comp["nbits"],
                                   int(cases[c, p]) & ((1 << comp["nbits"]) - 1))
        w[:, p] = rec
        for c, comp in enumerate(comps):
            code = orc.lib().orc_get_bits(orc._ptr(rec), comp["bit_offset"], comp["nbits"])
            v = orc.lib().orc_sext(code, comp["nbits"]) if comp["is_signed"] else code
            assert v == cases[c, p]
    # the 8-bit field straddles: 4 bits in word 0, 4 in word 1
    assert ((int(w[1, 0]) & 0xF) << 4 | (int(w[0, 0]) >> 60)) == (-(2 ** 7)) & 0xFF


def test_bit_struct_load_store(orc):
    """test_bit_struct.py:38-77: ci13, cu14, ci5 in bit_struct(32)."""
    comps = [cint(13, True, 0), cint(14, False, 13), cint(5, True, 27)]
    cases = np.array([[2 ** 12 - 1, 2 ** 14 - 1, -(2 ** 3)], [2 ** 11 - 1, 2 ** 13 - 1, -(2 ** 2)],
                      [0, 0, 0], [123, 4567, 8], [10, 31, 11]], dtype=np.float32).T
    out, _ = roundtrip(orc, comps, 1, cases)
    # 8 does not fit ci5 (max 15? ci5 is -16..15): the reference case [123, 4567, 8] fits
    assert np.array_equal(out, cases)


@pytest.mark.parametrize("phys,bits,case", [
    (8, [3, 3, 2], [2 ** 2 - 1, 2 ** 3 - 1, -2 ** 1]),
    (16, [4, 7, 5], [2 ** 3 - 1, 2 ** 7 - 1, -2 ** 4]),
    (32, [17, 11, 4], [2 ** 16 - 1, 2 ** 10 - 1, -2 ** 3]),
    (64, [32, 23, 9], [2 ** 31 - 1, 2 ** 23 - 1, -2 ** 8]),
    (32, [7, 12, 13], [2 ** 6 - 1, 2 ** 12 - 1, -2 ** 12]),
    (64, [18, 22, 24], [2 ** 17 - 1, 2 ** 22 - 1, -2 ** 23]),
    (16, [5, 5, 6], [15, 5, 20]),
    (32, [10, 10, 12], [11, 19, 2020]),
])
def test_bit_struct_widths(orc, phys, bits, case):
    """test_bit_struct.py:97-141: signed/unsigned/signed members of one bit_struct."""
    offs = [0, bits[0], bits[0] + bits[1]]
    assert offs[2] + bits[2] <= phys
    comps = [cint(bits[0], True, offs[0]), cint(bits[1], False, offs[1]), cint(bits[2], True, offs[2])]
    rec = np.zeros(1, dtype=np.uint64)
    for c, comp in enumerate(comps):
        orc.lib().orc_put_bits(orc._ptr(rec), comp["bit_offset"], comp["nbits"],
                               case[c] & ((1 << comp["nbits"]) - 1))
    for c, comp in enumerate(comps):
        code = orc.lib().orc_get_bits(orc._ptr(rec), comp["bit_offset"], comp["nbits"])
        v = orc.lib().orc_sext(code, comp["nbits"]) if comp["is_signed"] else code
        assert v == case[c]


def test_custom_int_implicit_cast(orc):
    """test_custom_int.py:4-16: storing 10.3 into ci13 gives 10."""
    out, _ = roundtrip(orc, [cint(13, True, 0)], 1, [[10.3]])
    assert out[0, 0] == 10


# --- test_custom_float.py / test_fxp_offset.py --------------------------------

def test_custom_float_fixed32(orc):
    """test_custom_float.py:8-26: fixed32 range 2 in a bit_pack; 0.7 + 0.4, 0.64, 0.66."""
    comps = [fixed(32, 2.0, off=0)]
    out, _ = roundtrip(orc, comps, 1, [[0.7]])
    out2, _ = roundtrip(orc, comps, 1, [[np.float32(out[0, 0]) + np.float32(0.4)]])
    assert out2[0, 0] == approx(1.1)
    for v in (0.64, 0.66):
        assert roundtrip(orc, comps, 1, [[v]])[0][0, 0] == approx(v)


@pytest.mark.parametrize("use_bit_pack", [True, False])
def test_custom_matrix_rotation(orc, use_bit_pack):
    """test_custom_float.py:29-59: five 18-degree rotations of a quantized 2x2 matrix."""
    if use_bit_pack:
        comps = [fixed(28, 1.2, off=28 * i) for i in range(4)]
        nw = 2
    else:
        comps = [fixed(16, 1.2, off=0), fixed(16, 1.2, off=16),
                 fixed(16, 1.2, off=32), fixed(16, 1.2, off=48)]
        nw = 1
    w = orc.pack(comps, nw, np.array([[1.0], [0.0], [0.0], [1.0]], dtype=np.float32))
    ang = math.pi / 10
    c, s = np.float32(math.cos(ang)), np.float32(math.sin(ang))
    R = np.array([[c, s], [-s, c]], dtype=np.float32)
    for _ in range(5):
        x = orc.unpack(comps, nw, w)[:, 0].reshape(2, 2)
        y = np.zeros((2, 2), dtype=np.float32)
        for i in range(2):
            for j in range(2):
                y[i, j] = x[i, 0] * R[0, j] + x[i, 1] * R[1, j]
        w = orc.pack(comps, nw, y.reshape(4, 1))
    x = orc.unpack(comps, nw, w)[:, 0].reshape(2, 2)
    assert x == approx(np.array([[0, 1], [-1, 0]]), abs=1e-4)


def test_custom_float_implicit_cast_scale(orc):
    """test_custom_float.py:62-75: unsigned ci13 with scale 0.1 stores 10 as 10.0."""
    comp = dict(fixed(13, 1.0, signed=False), scale=0.1)
    assert roundtrip(orc, [comp], 1, [[10.0]])[0][0, 0] == approx(10.0)


def test_cache_read_only(orc):
    """test_custom_float.py:78-97: unsigned ci15, scale 0.1: 0.7 and 1.2 read back equal in f32."""
    comp = dict(fixed(15, 1.0, signed=False), scale=0.1)
    for v in (0.7, 1.2):
        assert roundtrip(orc, [comp], 1, [[v]])[0][0, 0] == np.float32(v)


def test_fxp_offset(orc):
    """test_fxp_offset.py:9-29: fixed29 range 2 offset 1."""
    comp = fixed(29, 2.0, offset=1.0)
    out = roundtrip(orc, [comp], 1, [[0.7]])[0][0, 0]
    assert roundtrip(orc, [comp], 1, [[out + np.float32(0.4)]])[0][0, 0] == approx(1.1, 1e-4)
    for v in (0.64, 0.66, 2.8):
        assert roundtrip(orc, [comp], 1, [[v]])[0][0, 0] == approx(v, 1e-4)


# --- test_custom_float_exponents.py -------------------------------------------

def _f(orc, v, nb, signed, eb):
    e, d = orc.float_encode(v, nb, signed, eb)
    return orc.float_decode(e, d, nb, signed, eb)


def test_custom_float_unsigned(orc):
    """test_custom_float_exponents.py:8-28: cu13 digits, 6-bit exponent."""
    for v in [0, 1 / 1024, 1.75 / 1024, 0.25, 0.5, 0.75, 1, 2, 3, 4, 5, 6, 7, 128, 256, 512, 1024]:
        assert _f(orc, v, 13, False, 6) == v


def test_custom_float_signed(orc):
    """test_custom_float_exponents.py:31-58."""
    for v in [0, 0.125, 0.5, 2, 4, 6, 7, 8, 9]:
        assert _f(orc, v, 13, True, 6) == v
        assert _f(orc, -v, 13, True, 6) == -v
    for v in [1e-30, 1e-20, 1e-10, 1e-2]:
        assert _f(orc, v, 13, True, 6) == approx(v, abs=1e-5)
        assert _f(orc, -v, 13, True, 6) == approx(-v, abs=1e-5)


@pytest.mark.parametrize("digits_bits", [23, 24])
def test_custom_float_precision(orc, digits_bits):
    """test_custom_float_exponents.py:61-83."""
    for v in [np.float32(np.pi), np.float32(np.pi * (1 << 100))]:
        r = np.float32(_f(orc, float(v), digits_bits, True, 8))
        if digits_bits == 24:
            assert r == v
        else:
            assert r != v
            assert r == approx(v, rel=3e-7)


@pytest.mark.parametrize("signed", [True, False])
def test_custom_float_truncation(orc, signed):
    """test_custom_float_exponents.py:86-114: 2-bit digits, 5-bit exponent."""
    for v in [1, 1.5]:
        assert _f(orc, v, 2, signed, 5) == v
    assert _f(orc, 1.75, 2, signed, 5) == (2 if signed else 1.75)
    assert _f(orc, 1.625, 2, signed, 5) == (1.5 if signed else 1.75)


def test_custom_float_atomic_demotion(orc):
    """test_custom_float_exponents.py:117-139: x += 1 twice from 0 gives 2."""
    x = 0.0
    for _ in range(2):
        x = _f(orc, x + 1, 2, True, 5)
    assert x == 2


# --- test_custom_float_shared_exp.py ------------------------------------------

def _shared_pair(eb, signed_b=False):
    # exponent placed before its first user, members follow (snode_expr_utils.cpp:44-66)
    return [cfloat(10, False, eb, off=eb, exp_off=0, shared=True),
            cfloat(14, signed_b, eb, off=eb + 10, exp_off=0, shared=True)]


class _Pair:
    def __init__(self, orc, comps):
        self.orc, self.comps = orc, comps
        self.rec = np.zeros(1, dtype=np.uint64)

    def get(self):
        return self.orc.unpack(self.comps, 1, self.rec.reshape(1, 1))[:, 0]

    def set(self, c, v):
        self.orc.store_one(self.comps, self.rec, c, float(v))

    def set_both(self, a, b):
        self.rec[:] = self.orc.pack(self.comps, 1, np.array([[a], [b]], dtype=np.float32))[:, 0]


@pytest.mark.parametrize("eb", [5, 6, 7, 8])
def test_shared_exponents(orc, eb):
    """test_custom_float_shared_exp.py:7-75."""
    x = _Pair(orc, _shared_pair(eb))
    assert list(x.get()) == [0.0, 0.0]
    x.set(0, 10)
    assert list(x.get()) == [10.0, 0.0]
    x.set(0, 0)
    assert list(x.get()) == [0.0, 0.0]
    x.set_both(3.2, 0.25)
    a, b = x.get()
    assert a == approx(3.2, rel=1e-3) and b == approx(0.25, rel=2e-2)
    x.set(0, 0.27)
    a, b = x.get()
    assert a == approx(0.27, rel=1e-2) and b == approx(0.25, rel=2e-2)
    x.set(0, 100)
    a, b = x.get()
    assert a == approx(100, rel=1e-3) and b == approx(0.25, rel=1e-2)
    x.set(1, 0)
    a, b = x.get()
    assert a == approx(100, rel=1e-3) and b == 0
    x.set_both(0, 0)
    assert list(x.get()) == [0.0, 0.0]
    x.set_both(1000, 1e-6)
    assert list(x.get()) == [1000.0, 0.0]
    x.set_both(1000, 1000)
    assert list(x.get()) == [1000.0, 1000.0]
    x.set_both(1e-30, 1e-30)
    a, b = x.get()
    if eb == 8:
        assert a == approx(1e-30, 1e-3) and b == approx(1e-30, 1e-4)
    else:
        assert a == 0 and b == 0


@pytest.mark.parametrize("eb", [5, 6, 7, 8])
def test_shared_exponent_add(orc, eb):
    """test_custom_float_shared_exp.py:78-114."""
    x = _Pair(orc, _shared_pair(eb))
    x.set(0, 4)
    assert list(x.get()) == [4, 0]
    x.set(1, 3)
    assert list(x.get()) == [4, 3]
    x.set(1, x.get()[1] + 1)
    assert list(x.get()) == [4, 4]
    for i in range(100):
        x.set(0, x.get()[0] + 4)
        x.set(1, x.get()[1] + 1)
        assert list(x.get()) == [4 + (i + 1) * 4, 4 + (i + 1)]


@pytest.mark.parametrize("eb", [5, 6, 7, 8])
def test_shared_exponent_borrow(orc, eb):
    """test_custom_float_shared_exp.py:117-146."""
    x = _Pair(orc, _shared_pair(eb))
    x.set_both(0, 100)
    for i in range(100):
        assert list(x.get()) == [i, 100 - i]
        x.set(0, x.get()[0] + 1)
        x.set(1, x.get()[1] - 1)


@pytest.mark.parametrize("eb", [5, 6, 7, 8])
def test_shared_negative(orc, eb):
    """test_custom_float_shared_exp.py:149-170."""
    x = _Pair(orc, _shared_pair(eb, signed_b=True))
    x.set(0, 37)
    assert x.get()[0] == 37
    x.set(1, -123)
    assert x.get()[1] == -123


# --- test_custom_float_time_integration.py -------------------------------------

@pytest.mark.parametrize("variant", ["f32", "fixed", "float", "shared"])
def test_custom_float_time_integration(orc, variant):
    """test_custom_float_time_integration.py:9-63: midpoint rotation for one period."""
    if variant == "fixed":
        comps = [dict(fixed(16), scale=1 / 2 ** 14), dict(fixed(16, off=16), scale=1 / 2 ** 14)]
    elif variant == "float":
        comps = [cfloat(13, True, 6, off=6, exp_off=0), cfloat(13, True, 6, off=32 + 6, exp_off=32)]
    elif variant == "shared":
        comps = [cfloat(13, True, 6, off=6, exp_off=0, shared=True),
                 cfloat(13, True, 6, off=19, exp_off=0, shared=True)]
    else:
        comps = [dict(kind=orc.F32, nbits=32, is_signed=1, bit_offset=0, scale=1.0),
                 dict(kind=orc.F32, nbits=32, is_signed=1, bit_offset=32, scale=1.0)]
    w = orc.pack(comps, 1, np.array([[1.0], [0.0]], dtype=np.float32))
    num_steps = 1000
    dt = np.float32(math.pi * 2 / num_steps)
    half = np.float32(0.5)
    px, py = [], []

    def v_at(p):
        return np.array([-p[1], p[0]], dtype=np.float32)

    for _ in range(num_steps):
        x = orc.unpack(comps, 1, w)[:, 0]
        v_mid = v_at(x + half * dt * v_at(x))
        w = orc.pack(comps, 1, (x + v_mid * dt).reshape(2, 1).astype(np.float32))
        xn = orc.unpack(comps, 1, w)[:, 0]
        px.append(xn[0])
        py.append(xn[1])
    assert px[num_steps // 2 - 1] == approx(-1, abs=2e-2)
    assert py[num_steps // 2 - 1] == approx(0, abs=2e-2)
    assert px[-1] == approx(1, abs=2e-2)
    assert py[-1] == approx(0, abs=3e-2)


# --- test_dithering.py (invariant form) ---------------------------------------

def test_dithering_ratio(orc):
    """test_dithering.py:4-31: fixed5 range 2, f64 compute, dithered; 1.1 rounds up
    to 1.125 with probability 0.8 and down to 1.0 with 0.2, ratio 4 +- 0.1.

    The reference reaches 1.1 by forwarding the stored 0.7 unquantized to the
    following read; here 1.1 is encoded directly (one sample per slot)."""
    comp = fixed(5, 2.0, dither=True)
    n = 1000000
    vals = np.full((1, n), 1.1, dtype=np.float64)
    w = orc.pack([comp], 1, vals, f64=True, seed=1234, step=0)
    out = orc.unpack([comp], 1, w, f64=True)[0]
    up = int(np.sum(out == 1.125))
    down = int(np.sum(out == 1.0))
    assert up + down == n
    # 10^6 samples: the up-fraction has std 4e-4, the ratio std 0.01
    assert abs(up / n - 0.8) < 0.002
    assert abs(up / down - 4) < 0.1


def test_dither_noise_grid_uniformity_and_pairs(orc):
    """The codec's dithering noise (one hash per component pair, 16 bits per
    component): every value is (h + 1/2) 2^-16 with h in 0..65535; over many
    slots each component's values are uniform (mean 1/2, variance 1/12,
    every 1/16 bin near 1/16) and the two components of a pair are
    uncorrelated, as are neighbouring slots."""
    n = 200000
    u = np.array([[orc.dither_u(99, 7, p, c) for p in range(n)] for c in (4, 5)], dtype=np.float64)
    h = u * 65536.0 - 0.5
    assert np.all(h == np.round(h)) and h.min() >= 0 and h.max() <= 65535
    for row in u:
        assert abs(row.mean() - 0.5) < 3e-3
        assert abs(row.var() - 1.0 / 12.0) < 2e-3
        bins = np.bincount(np.minimum((row * 16).astype(int), 15), minlength=16) / n
        assert np.all(np.abs(bins - 1 / 16) < 4e-3)
    assert abs(np.corrcoef(u[0], u[1])[0, 1]) < 1e-2
    assert abs(np.corrcoef(u[0, :-1], u[0, 1:])[0, 1]) < 1e-2
    # other pairs and steps draw other values
    assert orc.dither_u(99, 7, 11, 6) != orc.dither_u(99, 7, 11, 4)
    assert orc.dither_u(99, 8, 11, 4) != orc.dither_u(99, 7, 11, 4)
