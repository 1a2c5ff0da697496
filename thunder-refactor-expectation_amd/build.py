"""Build recipe of the expectation-step library and of the CPU oracle.

`build_library()` compiles the HIP sources for gfx950 with hipcc into
`libthunder_expect.so` next to this file; `build_oracle()` compiles the C
oracle (test infrastructure) with gcc into `oracle/liboracle_thx.so`.
Both always recompile.  Run as `python build.py` from anywhere.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libthunder_expect.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle_thx.so")

HIP_SOURCES = ["thx_kernels.hip", "thx_pf.hip", "thx_cnn.hip", "thx_filter.hip",
               "thx_projectee.hip", "thx_images.hip", "thx_sigma.hip", "thx_insert.hip",
               "thx_scale.hip", "thx_noalloc.hip",
               "thx_api.hip"]
# host C++ built into the same library: the adapter that keeps the reference's
# Projector / Optimiser::expectation / ExpectLocal* call surfaces, and the
# .thu / MRC file formats
CXX_SOURCES = ["thx_adapter.cpp", "thx_io.cpp"]
ADAPTER_CHECK = os.path.join(ROOT, "tests", "cpp", "adapter_check")
# hipFFT: the two plain 3-D transforms of the projectee preparation and the
# batched 2-D transform of the particle images
HIP_LIBS = ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]

HIPCC_FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed FP32 VALU ops are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(PKG, "csrc", s) for s in HIP_SOURCES]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc")]
    objs = []
    for s in CXX_SOURCES:
        obj = os.path.join(PKG, "csrc", os.path.splitext(s)[0] + ".o")
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-c"] + inc +
             ["-o", obj, os.path.join(PKG, "csrc", s)])
        objs.append(obj)
    tmp = LIB + ".tmp"
    # "-x none": the objects after the HIP sources are linked, not compiled as HIP
    _run([hipcc] + HIPCC_FLAGS + inc + ["-o", tmp] + srcs + ["-x", "none"] + objs + HIP_LIBS)
    os.replace(tmp, LIB)
    build_adapter_check()
    return LIB


def build_adapter_check():
    """The C++ driver of the adapter test (links the library)."""
    tmp = ADAPTER_CHECK + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-o", tmp,
          os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp"), "-L", PKG, "-lthunder_expect",
          "-Wl,-rpath,$ORIGIN/../../thunder-refactor-expectation_amd"])
    os.replace(tmp, ADAPTER_CHECK)
    return ADAPTER_CHECK


def build_oracle():
    src = os.path.join(ROOT, "oracle", "thx_oracle.c")
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-fopenmp", "-mavx2", "-mfma", "-fPIC", "-shared",
          "-std=gnu11", "-o", tmp, src, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


if __name__ == "__main__":
    what = sys.argv[1:] or ["lib", "oracle"]
    if "oracle" in what:
        build_oracle()
    if "lib" in what:
        build_library()
