"""Run-time specialisation of the fused kernel to a record layout
(aq_jit_build / aq_sim_specialize), checked without a device: the generated
source compiles with hipRTC for gfx950 for the benchmark record, the 2D
elastic record and a record whose fields are out of component order; the
code object holds the kernel, uses no scratch memory and no packed-FP32
instructions; the cache returns the same object; plain storage is refused."""
import os
import re
import shutil
import subprocess

import pytest

import autoquantizer_amd as aq
from autoquantizer_amd import _lib, layouts

LLVM = "/opt/rocm/llvm/bin"


def fluid_cfg():
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 128, 2e-4, 400.0
    cfg.p_vol = (1 / 128 * 0.5) ** 3
    cfg.n_particles = 1 << 20
    cfg.compute_fp = "f32"
    return cfg


def elastic_cfg():
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu, cfg.sticky = 2, False, 128, 1e-4, 1e4, 0.2, True
    cfg.p_vol = (1 / 128 * 0.5) ** 2
    cfg.n_particles = 8192
    cfg.compute_fp = "f32"
    return cfg


def descs():
    _, t, g, sh = layouts.fluid_3d()
    out = {"fluid_3d": aq.MPMSolver.describe(fluid_cfg(), types=t, groups=g, shared=sh)}
    _, t2, g2, sh2 = layouts.elastic_2d()
    out["elastic_2d"] = aq.MPMSolver.describe(elastic_cfg(), types=t2, groups=g2, shared=sh2)
    order = ["J"] + [f"C{a}{b}" for a in range(3) for b in range(3)] + [f"v{a}" for a in range(3)] + \
            [f"x{a}" for a in range(3)]
    out["fluid_3d_reordered"] = aq.MPMSolver.describe(fluid_cfg(), types=t, groups=[("pack", 0, order)], shared=sh)
    return out


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    cache = str(tmp_path_factory.mktemp("jit"))
    return cache, {k: _lib.jit_build(d, cache) for k, d in descs().items()}


@pytest.mark.parametrize("kind", ["fluid_3d", "elastic_2d", "fluid_3d_reordered"])
def test_code_object(built, kind):
    cache, paths = built
    path = paths[kind]
    assert os.path.dirname(path) == cache and os.path.getsize(path) > 1000
    with open(path, "rb") as f:
        assert f.read(4) == b"\x7fELF"
    if not os.path.exists(os.path.join(LLVM, "llvm-readelf")):
        pytest.skip("llvm tools not installed")
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", path], capture_output=True,
                           text=True).stdout
    for name in ("aq_jit_tiled", "aq_jit_p2g", "aq_jit_g2p"):  # fused, P2G only, G2P only
        assert name in notes
    sizes = re.findall(r"\.private_segment_fixed_size:\s*(\d+)", notes)
    assert len(sizes) == 3 and all(v == "0" for v in sizes), sizes  # no register spills
    dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", path], capture_output=True, text=True).stdout
    assert "ds_add_u64" in dis
    for op in ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32"):
        assert op not in dis


def test_cache_returns_the_same_object(built):
    cache, paths = built
    stamp = os.path.getmtime(paths["fluid_3d"])
    _, t, g, sh = layouts.fluid_3d()
    again = _lib.jit_build(aq.MPMSolver.describe(fluid_cfg(), types=t, groups=g, shared=sh), cache)
    assert again == paths["fluid_3d"] and os.path.getmtime(again) == stamp
    # a different scheme (another bit width) is another object
    _, t2, g2, sh2 = layouts.fluid_3d(v_range=4.0)
    other = _lib.jit_build(aq.MPMSolver.describe(fluid_cfg(), types=t2, groups=g2, shared=sh2), cache)
    assert other != paths["fluid_3d"]


def test_plain_storage_is_refused(tmp_path):
    cfg = fluid_cfg()
    cfg.compute_fp = "f32"
    with pytest.raises(_lib.AQError):
        _lib.jit_build(aq.MPMSolver.describe(cfg), str(tmp_path))
