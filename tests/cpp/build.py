"""Builds tests/cpp/_build/facade_check: a C++ program that uses include/skwarka.hpp as the
reference's callers use its classes, linked against the in-tree libskwarka_rt.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "facade_check")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.join(ROOT, "skwarka_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROCM, "include"), os.path.join(HERE, "facade_check.cpp"), "-o", OUT + ".tmp",
           "-L" + libdir, "-lskwarka_rt", "-Wl,-rpath," + libdir, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("facade_check build failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


KD_OUT = os.path.join(HERE, "_build", "kd_select_check")


def build_kd_select():
    """tests/cpp/kd_select_check: the restated libstdc++ selection against libstdc++ (host only)."""
    os.makedirs(os.path.dirname(KD_OUT), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "kd_select_check.cpp"), "-o", KD_OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("kd_select_check build failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(KD_OUT + ".tmp", KD_OUT)
    return KD_OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
