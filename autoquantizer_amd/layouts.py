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
hand_scheme_2d: the reference's hand-written 2D elastic scheme
  (MPMSolver.init_hand_scheme, mpm_exps/engine/mpm_solver.py:325-362)
  x  2 x fixed20, unsigned, range 2 (or 1 with refined ranges)  bit_struct(64)
  v  2 x float(exp=6, frac=10), own exponents       bit_struct(32)
     (refined ranges: 2 x fixed16, range 18)
  F  4 x fixed16, range 4 (refined: the F ranges)    two bit_struct(32)
  C  unquantized f32                                 128 bits
  dithered fixed-point members when dither is set.
Custom floats with exponents inside a bit_pack, and a shared exponent group
spanning two words, are extensions: the reference allows exponents only in
a bit_struct of at most 64 bits (TS/taichi/codegen/codegen_llvm_quant.cpp:766-767).
"""
from .mpm import component_names
from .quant import PlainF32, quant


def hand_scheme_2d(dither=True, refined_ranges=None):
    """The record of init_hand_scheme (mpm_solver.py:325-362): names, types,
    groups, shared.  refined_ranges: None for the hand ranges, else the four
    F ranges (the refine_ranges branch, :343-358, which also narrows x to
    range 1 and makes v fixed16 of range 18)."""
    names = component_names(2, False)
    t = {}
    refined = refined_ranges is not None
    for a in range(2):
        t[f"x{a}"] = quant.fixed(20, range=1.0 if refined else 2.0, signed=False, use_dithering=dither)
        t[f"v{a}"] = (quant.fixed(16, range=18.0, use_dithering=dither) if refined
                      else quant.float(exp=6, frac=10, signed=True))
    for k, nm in enumerate(("F00", "F01", "F10", "F11")):
        r = float(refined_ranges[k]) if refined else 4.0
        t[nm] = quant.fixed(16, range=r, use_dithering=dither)
    for nm in ("C00", "C01", "C10", "C11"):
        t[nm] = PlainF32()
    groups = [("struct", 64, ["x0", "x1"]), ("struct", 32, ["v0", "v1"]),
              ("struct", 32, ["F00", "F01"]), ("struct", 32, ["F10", "F11"]),
              ("pack", 0, ["C00", "C01", "C10", "C11"])]
    return names, t, groups, ()


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


def fluid_3d_error_bounds(nsteps, n_grid, v_range=8.0):
    """Stated bound on the RMS error over particles, per field, of a run of
    nsteps substeps with the fluid_3d record against an FP64 run of the same
    scene.  Each store rounds a fixed-point member to a neighbouring code with
    dithering (error below one step q, mean zero, RMS at most q/2):
      x, J  integrate their own rounding errors like a random walk:
            q sqrt(nsteps);
      v     is gathered afresh from the grid every substep, so only the last
            rounding and the averaged noise of the neighbours remain: q_v;
      C     = 4 inv_dx sum_i w_i v_i (x_i - x_p)^T is the stencil derivative
            of the grid velocity, whose noise is the particles' v rounding
            averaged over about ten particles per node: inv_dx q_v / 4 (its
            own 10 digit bits add a relative 2^-10, far below).
    q: x 2/2^21 (range 1), v 2 v_range/2^16, J 1/2^16 (range 0.5)."""
    import math
    qx, qv, qj = 2.0 / 2 ** 21, 2.0 * v_range / 2 ** 16, 1.0 / 2 ** 16
    r = math.sqrt(max(nsteps, 1))
    return {"x": qx * r, "v": qv, "C": n_grid * qv / 4.0, "J": qj * r}


def fluid_3d(dither=True, v_range=8.0, x_range=1.0):
    names = component_names(3, True)
    t = {}
    for a in range(3):
        t[f"x{a}"] = quant.fixed(21, signed=True, range=x_range, use_dithering=dither)
        t[f"v{a}"] = quant.fixed(16, signed=True, range=v_range, use_dithering=dither)
        for b in range(3):
            t[f"C{a}{b}"] = quant.float(exp=8, frac=10)
    t["J"] = quant.fixed(16, signed=True, range=0.5, offset=1.0, use_dithering=dither)
    groups = [("pack", 0, names)]
    return names, t, groups, ("C",)
