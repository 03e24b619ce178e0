"""Build libakx.so (HIP, gfx950) in-tree with hipcc, then the C++ API over it
(libakantu_amd.so, g++) and its example drivers (facade/bin/).

Flags:
  -ffp-contract=off       the FP64 path keeps the reference's operation order
                          (no fused multiply-add; the FP32 kernels opt back in
                          locally with `#pragma clang fp contract(fast)`);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libakx.so")
ARCH = os.environ.get("AKX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["akx_kernels.hip", "akx_surrogate.hip", "akx_fe.hip", "akx_noalloc.hip",
           "akx_api.cpp"]
HEADERS = ["akx_device.hpp", "akx_internal.hpp", "akx_mlp.hpp"]
FACADE = os.path.join(HERE, "facade")
FACADE_LIB = os.path.join(HERE, "libakantu_amd.so")
EXAMPLES = ["explicit_dynamic", "patch_test_explicit", "fe_engine_operators"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def common_flags():
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=%s" % ARCH, "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]


def build(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + common_flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "--offload-arch=%s" % ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_facade(verbose)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_facade(verbose: bool = False) -> str:
    """libakantu_amd.so (the akantu:: C++ API, host code only) linked against
    libakx.so, and the example drivers in facade/bin/ (run-paths point back at
    the package directory, so the tree can be moved as a whole)."""
    cxx = os.environ.get("CXX", "g++")
    inc = ["-I", os.path.join(FACADE, "include"), "-I", os.path.join(ROOT, "include")]
    flags = ["-O2", "-std=c++17", "-pthread", "-Wall", "-Wextra"]
    tmp = FACADE_LIB + ".tmp"
    _run([cxx] + flags + ["-fPIC", "-shared"] + inc +
         [os.path.join(FACADE, "src", "akantu.cc"), "-o", tmp, "-L", HERE, "-lakx",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, FACADE_LIB)
    bdir = os.path.join(FACADE, "bin")
    os.makedirs(bdir, exist_ok=True)
    for ex in EXAMPLES:
        out = os.path.join(bdir, ex)
        _run([cxx] + flags + inc + [os.path.join(FACADE, "examples", ex + ".cc"), "-o",
                                    out + ".tmp", "-L", HERE, "-lakantu_amd", "-lakx",
                                    "-Wl,-rpath,$ORIGIN/../.."], verbose)
        os.replace(out + ".tmp", out)
    return FACADE_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
