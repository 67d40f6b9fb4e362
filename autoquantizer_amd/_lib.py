"""ctypes binding of libaqmpm.so (include/aq_mpm.h).

The library is the only compute path: if it is missing or cannot load, every
entry point raises.  There is no CPU fallback.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libaqmpm.so")
# kernel sources read by the run-time specialisation (aq_sim_specialize)
CSRC_DIR = os.path.join(_PKG, "csrc")
INCLUDE_DIR = os.path.join(os.path.dirname(_PKG), "include")


def jit_cache_dir(path=None):
    """Directory of cached specialised code objects (created with mode 0700
    if missing): path, else $AQ_JIT_CACHE, else aqmpm_jit_<uid> in the
    system temp dir.  The library loads and runs the code objects it finds
    there, so a directory owned by another user or writable by others is
    refused."""
    import stat
    import tempfile
    d = path or os.environ.get("AQ_JIT_CACHE") or os.path.join(tempfile.gettempdir(), f"aqmpm_jit_{os.getuid()}")
    os.makedirs(d, mode=0o700, exist_ok=True)
    st = os.stat(d)
    if st.st_uid != os.geteuid():
        raise AQError(f"JIT cache directory {d} is not owned by the current user")
    if st.st_mode & (stat.S_IWGRP | stat.S_IWOTH):
        raise AQError(f"JIT cache directory {d} is writable by other users")
    return d


def jit_build(desc, cache_dir=None):
    """Compile (or find cached) the specialised fused kernel for an
    aq_sim_desc without a device; returns the code object's path."""
    buf = ctypes.create_string_buffer(4096)
    check(lib().aq_jit_build(ctypes.byref(desc), CSRC_DIR.encode(), INCLUDE_DIR.encode(),
                             jit_cache_dir(cache_dir).encode(), buf, len(buf)))
    return buf.value.decode()

AQ_MAX_COMPS = 32
AQ_MAX_WORDS = 16

AQ_FIXED, AQ_INT, AQ_FLOAT, AQ_F32 = 0, 1, 2, 3
AQ_BIT_STRUCT, AQ_BIT_PACK = 0, 1
AQ_FLUID, AQ_ELASTIC = 0, 1
AQ_SEPARATING, AQ_STICKY = 0, 1
AQ_STORE_QUANT, AQ_STORE_F32, AQ_STORE_F64 = 0, 1, 2


class Component(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("num_bits", ctypes.c_int32),
                ("is_signed", ctypes.c_int32), ("bit_offset", ctypes.c_int32),
                ("exp_bits", ctypes.c_int32), ("exp_offset", ctypes.c_int32),
                ("shared", ctypes.c_int32), ("use_dithering", ctypes.c_int32),
                ("scale", ctypes.c_double), ("offset", ctypes.c_double)]


class Place(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("num_bits", ctypes.c_int32),
                ("is_signed", ctypes.c_int32), ("exp_bits", ctypes.c_int32),
                ("use_dithering", ctypes.c_int32), ("container", ctypes.c_int32),
                ("shared_group", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("scale", ctypes.c_double), ("offset", ctypes.c_double)]


class Container(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("num_bits", ctypes.c_int32)]


class SimDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("material", ctypes.c_int32),
                ("n_grid", ctypes.c_int32), ("boundary", ctypes.c_int32),
                ("bound", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("ncomp", ctypes.c_int32), ("nwords", ctypes.c_int32),
                ("capacity", ctypes.c_int64),
                ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double),
                ("dt", ctypes.c_double), ("p_vol", ctypes.c_double),
                ("p_mass", ctypes.c_double), ("gravity", ctypes.c_double),
                ("la", ctypes.c_double), ("mu_0", ctypes.c_double),
                ("la_0", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("comps", Component * AQ_MAX_COMPS),
                ("axis_bounds", ctypes.c_int32), ("lb", ctypes.c_int32 * 3), ("ub", ctypes.c_int32 * 3),
                ("bound_clip", ctypes.c_int32), ("compute_f64", ctypes.c_int32)]


class Collider(ctypes.Structure):
    _fields_ = [("sdf", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("size", ctypes.c_int32 * 3), ("offset", ctypes.c_int32 * 3), ("center", ctypes.c_int32 * 3),
                ("lag", ctypes.c_int32), ("velocity_from_rotated", ctypes.c_int32), ("first_step", ctypes.c_uint32),
                ("omega", ctypes.c_double), ("height_rate", ctypes.c_double)]


class AQError(RuntimeError):
    pass


AQ_MLP_IN, AQ_MLP_HIDDEN, AQ_MLP_OUT, AQ_MLP_PARAMS = 24, 64, 4, 10180  # include/aq_mpm.h
AQ_MLP_NOALLOC_MAX_SAMPLES = 2048

EXPORTS = [
    "aq_last_error", "aq_version", "aq_layout_plan", "aq_pack", "aq_unpack",
    "aq_store_members", "aq_sim_workspace_bytes", "aq_sim_create", "aq_sim_destroy",
    "aq_substep", "aq_p2g", "aq_grid_op", "aq_g2p", "aq_advance", "aq_sort",
    "aq_update_limits", "aq_advance_limits", "aq_energy", "aq_sim_grid", "aq_advance_timed",
    "aq_adj_workspace_bytes", "aq_adj_substep", "aq_noise_add", "aq_noise_dot", "aq_codec_selftest", "aq_mlp_forward",
    "aq_mlp_backward", "aq_adam_step", "aq_sim_specialize", "aq_jit_build", "aq_mlp_backward_noalloc",
    "aq_sim_count_wraps", "aq_sim_pack", "aq_sim_unpack", "aq_group_workspace_bytes", "aq_group_create", "aq_group_advance", "aq_group_destroy",
    "aq_sim_set_collider",
]

_lib = None


def lib():
    """Load libaqmpm.so; raises AQError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AQError(f"{LIB_PATH} is not built; run `python -m autoquantizer_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, i64, u32, u64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                             ctypes.c_uint32, ctypes.c_uint64)
    L.aq_last_error.restype = ctypes.c_char_p
    L.aq_last_error.argtypes = []
    L.aq_version.restype = ctypes.c_char_p
    L.aq_version.argtypes = []
    L.aq_layout_plan.argtypes = [P, i32, P, i32, P, P]
    L.aq_pack.argtypes = [P, i32, i32, P, P, u64, u32, P, i64, P]
    L.aq_unpack.argtypes = [P, i32, i32, P, P, i64, P]
    L.aq_store_members.argtypes = [P, i32, i32, P, i64, u32, P, P, i64, u64, u32, P]
    L.aq_sim_workspace_bytes.restype = i64
    L.aq_sim_workspace_bytes.argtypes = [P]
    L.aq_sim_create.argtypes = [P, P, i64, ctypes.POINTER(P)]
    L.aq_sim_destroy.argtypes = [P]
    L.aq_substep.argtypes = [P, P, i64, P, u32, P]
    L.aq_p2g.argtypes = [P, P, i64, P]
    L.aq_grid_op.argtypes = [P, u32, P]
    L.aq_sim_set_collider.argtypes = [P, P]
    L.aq_g2p.argtypes = [P, P, i64, P, u32, P]
    L.aq_advance.argtypes = [P, P, i64, i32, u32, i32, P, P]
    L.aq_sort.argtypes = [P, P, i64, P, P]
    L.aq_update_limits.argtypes = [P, P, i64, P, P]
    L.aq_advance_limits.argtypes = [P, P, i64, i32, u32, i32, P, P, P]
    L.aq_energy.argtypes = [P, P, i64, P, P]
    L.aq_sim_grid.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(i64)]
    L.aq_advance_timed.argtypes = [P, P, i64, i32, u32, i32, P, P, P, P, P]
    L.aq_adj_workspace_bytes.restype = i64
    L.aq_adj_workspace_bytes.argtypes = [P]
    L.aq_adj_substep.argtypes = [P, u32, P, i64, P, P, P, P, i64, P]
    L.aq_sim_specialize.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
    L.aq_jit_build.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i64]
    L.aq_noise_add.argtypes = [P, P, i64, P, P, u64, u32, P]
    L.aq_noise_dot.argtypes = [P, P, i64, P, u64, u32, P, P]
    L.aq_codec_selftest.argtypes = [P, P]
    L.aq_mlp_forward.argtypes = [P, P, i32, P, P, P]
    f32 = ctypes.c_float
    L.aq_mlp_backward.argtypes = [P, P, P, P, P, i32, f32, P, P, P, P]
    L.aq_sim_count_wraps.argtypes = [P, P]
    L.aq_sim_pack.argtypes = [P, P, i64, u32, P, P]
    L.aq_sim_unpack.argtypes = [P, P, i64, P, P]
    L.aq_mlp_backward_noalloc.argtypes = [P, P, P, P, P, i32, f32, P, P, P]
    L.aq_adam_step.argtypes = [P, P, P, P, i32, i32, f32, f32, f32, f32, P]
    L.aq_group_workspace_bytes.restype = i64
    L.aq_group_workspace_bytes.argtypes = [i32]
    L.aq_group_create.argtypes = [P, P, P, P, i32, P, i64, ctypes.POINTER(P)]
    L.aq_group_advance.argtypes = [P, i32, u32, i32, i32, P]
    L.aq_group_destroy.argtypes = [P]
    for name in EXPORTS:
        if name not in ("aq_last_error", "aq_version", "aq_sim_workspace_bytes", "aq_adj_workspace_bytes",
                        "aq_group_workspace_bytes"):
            getattr(L, name).restype = i32
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise AQError(f"libaqmpm error {rc}: {lib().aq_last_error().decode()}")
    return rc
