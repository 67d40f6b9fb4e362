"""CPU tests of the C ABI: the library loads, exports every function declared
in include/aq_mpm.h, and the host-side layout planner follows the
reference's placement rules (TS/taichi/struct/struct_llvm.cpp:71-116,
TS/taichi/program/snode_expr_utils.cpp:44-89).  No GPU is touched."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "aq_mpm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aq_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    from autoquantizer_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) == set(names)
    assert b"gfx950" in L.aq_version()


def q():
    import autoquantizer_amd as aq
    return aq


def test_plan_bit_pack_assign_layout():
    """test_bit_pack_assign.py:12-17: fixed21 x3 + int15 in a bit_pack: offsets
    0, 21, 42, 63 (the int straddles into word 1); two words."""
    aq = q()
    from autoquantizer_amd import _lib
    t = aq.quant.fixed(frac=21, range=2)
    comps, nw = aq.plan([(_lib.AQ_BIT_PACK, 0)],
                        [dict(t.place_record(), container=0, shared_group=-1)] * 3 +
                        [dict(kind=_lib.AQ_INT, num_bits=15, is_signed=0, exp_bits=0, use_dithering=0,
                              scale=1.0, offset=0.0, container=0, shared_group=-1)])
    assert [c["bit_offset"] for c in comps] == [0, 21, 42, 63]
    assert nw == 2
    assert comps[0]["scale"] == 2 / 2 ** 20


def test_plan_shared_exponent_first():
    """Shared exponent placed before its first user; members follow."""
    aq = q()
    from autoquantizer_amd import _lib
    ft = aq.quant.float(exp=6, frac=13)
    recs = [dict(ft.place_record(), container=0, shared_group=0) for _ in range(4)]
    comps, nw = aq.plan([(_lib.AQ_BIT_STRUCT, 64)], recs)
    assert [c["exp_offset"] for c in comps] == [0] * 4
    assert [c["bit_offset"] for c in comps] == [6, 19, 32, 45]
    assert all(c["shared"] == 1 for c in comps)
    assert nw == 1


def test_plan_non_shared_floats_own_exponent():
    aq = q()
    from autoquantizer_amd import _lib
    ft = aq.quant.float(exp=6, frac=13)
    recs = [dict(ft.place_record(), container=0, shared_group=-1) for _ in range(2)]
    comps, nw = aq.plan([(_lib.AQ_BIT_STRUCT, 64)], recs)
    assert [(c["exp_offset"], c["bit_offset"]) for c in comps] == [(0, 6), (19, 25)]
    assert all(c["shared"] == 0 for c in comps)


def test_plan_bit_struct_overflow_rejected():
    aq = q()
    from autoquantizer_amd import _lib
    t = aq.quant.fixed(frac=20)
    with pytest.raises(_lib.AQError):
        aq.plan([(_lib.AQ_BIT_STRUCT, 32)], [dict(t.place_record(), container=0, shared_group=-1)] * 2)


def test_plan_containers_align():
    """bit_struct(32) then bit_struct(64): the 64-bit container starts on a word."""
    aq = q()
    from autoquantizer_amd import _lib
    t = aq.quant.fixed(frac=16)
    recs = [dict(t.place_record(), container=0, shared_group=-1),
            dict(t.place_record(), container=1, shared_group=-1)]
    comps, nw = aq.plan([(_lib.AQ_BIT_STRUCT, 32), (_lib.AQ_BIT_STRUCT, 64)], recs)
    assert [c["bit_offset"] for c in comps] == [0, 64]
    assert nw == 2


def test_config_layouts():
    """The two benchmark layouts: 2D elastic 24 B, 3D fluid 32 B per particle."""
    from autoquantizer_amd.mpm import component_names, plan_state
    from autoquantizer_amd import layouts
    names, types, groups, shared = layouts.elastic_2d()
    comps, nw = plan_state(names, types, groups, shared)
    assert nw == 3
    names, types, groups, shared = layouts.fluid_3d()
    comps, nw = plan_state(names, types, groups, shared)
    assert nw == 4
    assert comps[0]["bit_offset"] == 0 and comps[2]["bit_offset"] == 42


def test_sim_desc_validation():
    """Bad descriptors are rejected on the host with a message."""
    import ctypes
    from autoquantizer_amd import _lib
    L = _lib.lib()
    d = _lib.SimDesc()
    d.dim, d.material, d.n_grid, d.ncomp, d.storage = 3, _lib.AQ_ELASTIC, 64, 16, _lib.AQ_STORE_F32
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0  # 3D elastic carries 24 components
    assert b"ncomp" in L.aq_last_error()
    d.ncomp = 24
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) > 0
    d.material, d.ncomp = _lib.AQ_FLUID, 16
    d.n_grid = 66
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0
    d.n_grid, d.capacity = 64, 1000
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) > 64 ** 3 * 16 * 2
    # f64 compute belongs to quantized storage in 2D
    d.compute_f64 = 1
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0
    assert b"compute_f64" in L.aq_last_error()
    d.storage, d.nwords = _lib.AQ_STORE_QUANT, 4
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0  # 3D
    d2 = _lib.SimDesc()
    d2.dim, d2.material, d2.n_grid, d2.ncomp, d2.storage, d2.nwords = 2, _lib.AQ_ELASTIC, 64, 12, _lib.AQ_STORE_QUANT, 3
    d2.capacity, d2.compute_f64 = 1000, 1
    w64 = L.aq_sim_workspace_bytes(ctypes.byref(d2))
    d2.compute_f64 = 0
    w32 = L.aq_sim_workspace_bytes(ctypes.byref(d2))
    assert w64 > w32 > 0  # double grids and tiles
    d.storage, d.nwords, d.compute_f64 = _lib.AQ_STORE_F32, 0, 0
    # creation also needs a positive particle mass (checked before the workspace)
    h = ctypes.c_void_p()
    d.p_mass = 0.0
    assert L.aq_sim_create(ctypes.byref(d), None, 0, ctypes.byref(h)) < 0
    assert b"particle mass" in L.aq_last_error()


def test_header_constants_match_binding():
    from autoquantizer_amd import _lib
    txt = open(os.path.join(ROOT, "include", "aq_mpm.h")).read()
    defs = dict(re.findall(r"#define (AQ_[A-Z_0-9]+) \(?(-?\d+)\)?", txt))
    for name in ("AQ_MAX_COMPS", "AQ_MAX_WORDS", "AQ_MLP_IN", "AQ_MLP_HIDDEN", "AQ_MLP_OUT", "AQ_MLP_PARAMS",
                 "AQ_MLP_NOALLOC_MAX_SAMPLES"):
        assert int(defs[name]) == getattr(_lib, name), name
    h, i, o = _lib.AQ_MLP_HIDDEN, _lib.AQ_MLP_IN, _lib.AQ_MLP_OUT
    assert _lib.AQ_MLP_PARAMS == h * i + h + 2 * (h * h + h) + o * h + o


def test_noalloc_backward_rejects_before_device_work():
    """aq_mlp_backward_noalloc checks its sample count against the static
    buffer (and its other arguments) before it touches a device."""
    from autoquantizer_amd import _lib
    L = _lib.lib()
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails its checks first
    n = _lib.AQ_MLP_NOALLOC_MAX_SAMPLES + 1
    assert L.aq_mlp_backward_noalloc(p, p, p, p, p, n, 0.0, p, None, None) == -3
    assert b"static delta buffer" in L.aq_last_error()
    assert L.aq_mlp_backward_noalloc(p, p, p, p, p, 0, 0.0, p, None, None) == -1
    assert L.aq_mlp_backward_noalloc(p, p, p, p, p, 4, -1.0, p, None, None) == -1
    assert L.aq_mlp_backward_noalloc(p, p, p, p, p, 4, 0.0, None, None, None) == -1


def test_group_rejects_bad_arguments_on_the_host():
    """aq_group_* check their arguments before any device work."""
    from autoquantizer_amd import _lib
    L = _lib.lib()
    assert L.aq_group_workspace_bytes(0) < 0
    nb = L.aq_group_workspace_bytes(32)
    assert nb >= 32 * 64 and nb % 256 == 0
    V = ctypes.c_void_p * 1
    sims, states = V(None), V(0x1000)
    ns = (ctypes.c_int64 * 1)(10)
    h = ctypes.c_void_p()
    ws = ctypes.c_void_p(0x100000)  # never dereferenced: each call fails its checks first
    assert L.aq_group_create(sims, states, ns, None, 1, ws, nb, None) == -1
    assert L.aq_group_create(sims, states, ns, None, 0, ws, nb, ctypes.byref(h)) == -1
    assert L.aq_group_create(sims, states, ns, None, 1, ws, nb, ctypes.byref(h)) == -1
    assert b"null simulation" in L.aq_last_error()
    assert L.aq_group_create(sims, states, ns, None, 1, ctypes.c_void_p(0x100010), nb, ctypes.byref(h)) == -1
    assert b"aligned" in L.aq_last_error()
    assert L.aq_group_create(sims, states, ns, None, 1, ws, 16, ctypes.byref(h)) == -3
    assert L.aq_group_advance(None, 4, 0, 8, 0, None) == -1


def test_collider_and_walls_rejected_on_the_host():
    """aq_sim_set_collider and the per-axis walls check their arguments
    before any device work."""
    from autoquantizer_amd import _lib
    L = _lib.lib()
    c = _lib.Collider()
    assert L.aq_sim_set_collider(None, ctypes.byref(c)) == -1
    d = _lib.SimDesc()
    d.dim, d.material, d.n_grid, d.ncomp, d.storage, d.bound = 3, _lib.AQ_FLUID, 32, 16, _lib.AQ_STORE_F32, 3
    d.capacity = 100
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) > 0
    d.axis_bounds = 1
    d.lb[0], d.lb[1], d.lb[2] = 3, 3, 3
    d.ub[0], d.ub[1], d.ub[2] = 29, 13, 33
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0
    assert b"per-axis walls" in L.aq_last_error()
    d.ub[2] = 29
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) > 0
    d.bound_clip = 2
    assert L.aq_sim_workspace_bytes(ctypes.byref(d)) < 0
