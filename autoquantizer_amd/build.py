"""Builds libaqmpm.so (HIP, gfx950 only) in-tree.

    python -m autoquantizer_amd.build

Each source is compiled to an object in parallel, then the objects are
linked into the shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libaqmpm.so")
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("aq_mpm.hip", "aq_particles.hip", "aq_fused.hip", "aq_group.hip",
                                                  "aq_adjoint.hip", "aq_mlp.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OBJ_DIR = os.path.join(PKG, "_build")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every float operation of the codec and the physics is rounded on its own
    "-ffp-contract=off",
    # no packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
LINK = [
    "--offload-arch=gfx950", "-shared", "-fPIC",
    # hipRTC: run-time specialisation of the fused kernel to a record layout
    "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib",
]


def _compile(src, verbose):
    obj = os.path.join(OBJ_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
                             "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    cmd = [HIPCC] + LINK + ["-o", LIB + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
