"""Quantized types: the drop-in for ti.quant / ti.type_factory.

Mirrors TS/python/taichi/lang/quant_impl.py:13-73 and
TS/python/taichi/lang/type_factory_impl.py:27-54 (TS/ = taichi_submission/):
same names, argument meaning and defaults.  Types are plain value objects;
the bit layout happens at placement (autoquantizer_amd.fields).
"""
from dataclasses import dataclass
from typing import Optional

from . import _lib

f32 = "f32"
f64 = "f64"
i32 = "i32"


@dataclass(frozen=True)
class CustomIntType:
    """ti.quant.int (quant_impl.py:13-27): `bits` wide, signed or not."""
    bits: int
    signed: bool = False
    compute: Optional[str] = None

    def __post_init__(self):
        if not 1 <= self.bits <= 32:
            raise ValueError("custom int bits must be in [1, 32]")


@dataclass(frozen=True)
class CustomFloatType:
    """ti.type_factory.custom_float (type_factory_impl.py:27-54).

    Without exponent it is a fixed-point type: value = digits * scale + offset.
    With an exponent type it is a floating-point type (scale must be 1,
    compute type f32, no dithering: TS/taichi/ir/type.cpp:129-155)."""
    significand_type: CustomIntType
    exponent_type: Optional[CustomIntType] = None
    compute: Optional[str] = None
    scale: float = 1.0
    offset: float = 0.0
    use_dithering: bool = False

    def __post_init__(self):
        if self.exponent_type is not None:
            if self.exponent_type.signed or self.exponent_type.bits > 8:
                raise ValueError("exponent must be an unsigned custom int of <= 8 bits")
            if self.significand_type.bits - int(self.significand_type.signed) > 23:
                raise ValueError("custom float digits exceed 23 bits")
            if self.use_dithering:
                raise ValueError("custom floats with exponents do not dither")
            if self.compute not in (None, f32):
                raise ValueError("custom floats with exponents compute in f32")
            if self.scale != 1.0:
                raise ValueError("custom floats with exponents have scale 1")

    @property
    def bits(self):
        return self.significand_type.bits

    @property
    def signed(self):
        return self.significand_type.signed

    def place_record(self):
        if self.exponent_type is None:
            return dict(kind=_lib.AQ_FIXED, num_bits=self.bits, is_signed=int(self.signed),
                        exp_bits=0, use_dithering=int(self.use_dithering),
                        scale=float(self.scale), offset=float(self.offset))
        return dict(kind=_lib.AQ_FLOAT, num_bits=self.bits, is_signed=int(self.signed),
                    exp_bits=self.exponent_type.bits, use_dithering=0, scale=1.0, offset=0.0)


@dataclass(frozen=True)
class PlainF32:
    """An unquantized 32-bit float member."""
    def place_record(self):
        return dict(kind=_lib.AQ_F32, num_bits=32, is_signed=1, exp_bits=0, use_dithering=0,
                    scale=1.0, offset=0.0)


def _int_record(t: CustomIntType):
    return dict(kind=_lib.AQ_INT, num_bits=t.bits, is_signed=int(t.signed), exp_bits=0,
                use_dithering=0, scale=1.0, offset=0.0)


def place_record(t):
    if isinstance(t, CustomIntType):
        return _int_record(t)
    if isinstance(t, (CustomFloatType, PlainF32)):
        return t.place_record()
    if t in (f32, "float32", float):
        return PlainF32().place_record()
    raise TypeError(f"{t!r} cannot be placed in a bit container")


class _Quant:
    """Generator of quantized types (ti.quant, quant_impl.py:6-73)."""

    @staticmethod
    def int(bits, signed=False, compute=None):
        return CustomIntType(int(bits), bool(signed), compute)

    @staticmethod
    def fixed(frac, signed=True, range=1.0, offset=0.0, compute=None, use_dithering=False):
        """Fixed point with `frac` bits including the sign bit; scale is
        range / 2**(frac-1) signed, range / 2**frac unsigned (quant_impl.py:42-45)."""
        frac = int(frac)
        digits = CustomIntType(frac, bool(signed), i32)
        scale = range / 2 ** (frac - 1) if signed else range / 2 ** frac
        return CustomFloatType(digits, None, compute, float(scale), float(offset), bool(use_dithering))

    @staticmethod
    def float(exp, frac, signed=True, compute=None):
        """Floating point with an unsigned `exp`-bit exponent and `frac` digit
        bits including the sign bit (quant_impl.py:53-73)."""
        return CustomFloatType(CustomIntType(int(frac), bool(signed), i32),
                               CustomIntType(int(exp), False, i32), compute)


quant = _Quant()


class _TypeFactory:
    """ti.type_factory (type_factory_impl.py:5-54)."""

    @staticmethod
    def custom_int(bits, signed=False, compute_type=None):
        return CustomIntType(int(bits), bool(signed), compute_type)

    @staticmethod
    def custom_float(significand_type, exponent_type=None, compute_type=None, scale=1.0,
                     offset=0.0, use_dithering=False):
        return CustomFloatType(significand_type, exponent_type, compute_type, float(scale),
                               float(offset), bool(use_dithering))


type_factory = _TypeFactory()
