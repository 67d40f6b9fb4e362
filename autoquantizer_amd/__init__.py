"""autoquantizer_amd: the quantized MLS-MPM substep of Hanke98/AutoQuantizer
on MI355X (gfx950), as hand-written HIP kernels behind a C ABI
(include/aq_mpm.h, libaqmpm.so) with a Python surface mirroring the
reference's ti.quant / bit_pack / MPMSolver API.
"""
from . import _lib
from .quant import quant, type_factory, f32, f64, i32
from .fields import (field, Vector, Matrix, root, bit_pack_assign, i, j, RecordLayout, plan)
from .mpm import MPMConfig, MPMSolver, component_names

__all__ = ["quant", "type_factory", "f32", "f64", "i32", "field", "Vector", "Matrix", "root",
           "bit_pack_assign", "i", "j", "RecordLayout", "plan", "MPMConfig", "MPMSolver",
           "component_names", "lib"]


def lib():
    """The loaded libaqmpm.so (raises if it is not built)."""
    return _lib.lib()
