"""Particle record schemes of the benchmark scenes.

elastic_2d: 2D fixed-corotated jelly, 24 B per particle (fp32 state: 48 B)
  x  2 x fixed16, unsigned, range 1        (32 bits)
  v  2 x fixed16, signed, range 8          (32 bits)   -> one bit_pack word
  C  4 x float(exp=6, frac=13), one shared exponent (58 bits) -> bit_struct(64)
  F  4 x float(exp=6, frac=13), one shared exponent (58 bits) -> bit_struct(64)
  no dithering.
fluid_3d: 3D weakly compressible fluid, 32 B per particle (fp32 state: 64 B)
  x  3 x fixed21, signed, range 1          (63 bits)
  v  3 x fixed16, signed, range 8          (48 bits)
  C  9 x float(exp=8, frac=10), one shared exponent (98 bits)
  J  fixed16, signed, range 0.5, offset 1  (16 bits)
  all in one bit_pack of 225 bits (4 words), dithered fixed-point members.
Custom floats with exponents inside a bit_pack, and a shared exponent group
spanning two words, are extensions: the reference allows exponents only in
a bit_struct of at most 64 bits (TS/taichi/codegen/codegen_llvm_quant.cpp:766-767).
"""
from .mpm import component_names
from .quant import quant


def elastic_2d(dither=False):
    names = component_names(2, False)
    t = {}
    for a in range(2):
        t[f"x{a}"] = quant.fixed(16, signed=False, range=1.0, use_dithering=dither)
        t[f"v{a}"] = quant.fixed(16, signed=True, range=8.0, use_dithering=dither)
    for a in range(2):
        for b in range(2):
            t[f"C{a}{b}"] = quant.float(exp=6, frac=13)
            t[f"F{a}{b}"] = quant.float(exp=6, frac=13)
    groups = [("pack", 0, ["x0", "x1", "v0", "v1"]),
              ("struct", 64, ["C00", "C01", "C10", "C11"]),
              ("struct", 64, ["F00", "F01", "F10", "F11"])]
    return names, t, groups, ("C", "F")


def fluid_3d(dither=True, v_range=8.0):
    names = component_names(3, True)
    t = {}
    for a in range(3):
        t[f"x{a}"] = quant.fixed(21, signed=True, range=1.0, use_dithering=dither)
        t[f"v{a}"] = quant.fixed(16, signed=True, range=v_range, use_dithering=dither)
        for b in range(3):
            t[f"C{a}{b}"] = quant.float(exp=8, frac=10)
    t["J"] = quant.fixed(16, signed=True, range=0.5, offset=1.0, use_dithering=dither)
    groups = [("pack", 0, names)]
    return names, t, groups, ("C",)
