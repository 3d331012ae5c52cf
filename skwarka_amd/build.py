"""Builds libskwarka_rt.so (host builders + gfx950 kernels + C ABI) in-tree with hipcc.

Every HIP translation unit is compiled with -ffp-contract=off (the reference is an SSE2 build
without FMA, and bit-exact hit ids need the same rounding) and with packed FP32 instructions
disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libskwarka_rt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "abi.cpp",
    "abi_dl.cpp",
    "abi_batch.cpp",
    "abi_fg.cpp",
    "abi_pm.cpp",
    "abi_shoot.cpp",
    "abi_kd.cpp",
    "abi_noalloc.hip",
    "abi_tribuild.cpp",
    "io.cpp",
    "pbrt_import.cpp",
    "host/builders.cpp",
    "kernels/traverse.hip",
    "kernels/geometry.hip",
    "kernels/gather.hip",
    "kernels/dl.hip",
    "kernels/batch.hip",
    "kernels/sort.hip",
    "kernels/shoot.hip",
    "kernels/kd_build.hip",
    "kernels/tri_build.hip",
    "kernels/tri_build_wide.hip",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), res.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
