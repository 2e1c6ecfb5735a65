"""Build the in-tree HIP library ``libmercury_hip.so`` for gfx950.

hipcc cross-compiles without a GPU.  Flags that matter:
* ``-ffp-contract=off``: no FMA contraction, so the FP64 path performs the same
  rounded operations as the CPU restatement (bit-for-bit comparable);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: keep packed FP32
  arithmetic off (the host pass warns that it does not know the feature).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmercury_hip.so")
SOURCES = ["mercury_kernels.hip", "mercury_fused.hip", "mercury_resist.hip", "mercury_output.hip", "mercury_dec.hip", "mercury_singular_kernels.hip",
           "mercury_geometry.cpp", "mercury_io.cpp", "mercury_case.cpp", "mercury_patches.cpp", "mercury_singular.cpp", "mercury_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CHECKED_LIB = os.path.join(HERE, "libmercury_hip_checked.so")
# The checked build: the same device code, host code with bounds-checked
# standard containers (_GLIBCXX_ASSERTIONS) and undefined-behaviour checks that
# trap (UBSan in trap mode, no runtime library to load; vptr and function need
# one and are left out).  Host only: each -fsanitize flag goes to the host pass.
CHECK_FLAGS = ["-D_GLIBCXX_ASSERTIONS", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fsanitize-trap=undefined",
               "-Xarch_host", "-fno-sanitize=vptr,function"]


def build(verbose: bool = False, checked: bool = False) -> str:
    """Build the library; checked=True builds libmercury_hip_checked.so with the
    host-side checks of CHECK_FLAGS (select it with MERCURY_HIP_LIB)."""
    objs = []
    common = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
    if checked:
        common += CHECK_FLAGS
    out = CHECKED_LIB if checked else LIB
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ("_checked.o" if checked else ".o"))
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + common + lang + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


ROOT = os.path.dirname(HERE)
CALLER_SRC = os.path.join(ROOT, "tests", "cpp", "mercury_caller.cpp")
CALLER = os.path.join(ROOT, "tests", "cpp", "mercury_caller")


def build_cpp_caller(verbose: bool = False) -> str:
    """The C++ program of tests/cpp, written against the drop-in headers of
    include/mercury and linked with the in-tree library (host code only: g++)."""
    tmp = CALLER + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include", "mercury"),
           "-I", "/opt/rocm/include", CALLER_SRC, "-o", tmp, "-L", HERE, "-lmercury_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../mpcns_mercury_amd", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CALLER)
    return CALLER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, checked="--checked" in sys.argv))
