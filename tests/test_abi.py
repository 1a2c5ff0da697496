"""The C-ABI library loads and exports every symbol the header declares;
argument errors come back as status codes (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import thx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thunder_expect.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thx_\w+)\s*\(", text)))


def test_header_symbols_exported():
    lib = thx.load()
    names = declared()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", thx.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)
    # the binding covers exactly the declared entry points
    assert sorted(thx.exported_symbols()) == names


def test_argument_errors_are_status_codes():
    lib = thx.load()
    st = lib.thx_ExpectRotran(None, None, None, None, None, None, 1, 1, 8, 1)
    assert st == 1
    assert b"null" in lib.thx_last_error()
    st = lib.thx_ExpectProject(None, None, None, None, None, 1, 2, 1, 16, 1)
    assert st == 1
    h = C.c_void_p()
    st = lib.thx_ctx_create(-1, C.byref(h))
    assert st != 0 and not h
    assert lib.thx_ctx_destroy(None) == 0
    # producer side: a missing context is an argument error, before any device work
    st = lib.thx_ctx_set_particle_images(None, None, 4, 64, 30.0, 1, None, None, None)
    assert st == 1 and b"bad argument" in lib.thx_last_error()
    st = lib.thx_ctx_init_sigma(None, 31, 0, None, None, None)
    assert st == 1 and b"null context" in lib.thx_last_error()


def test_device_count_without_gpu_does_not_crash():
    n = C.c_int(-1)
    st = thx.load().thx_device_count(C.byref(n))
    assert st in (0, 2)
    assert n.value >= 0


def test_particle_arrays_must_match_the_pixel_set():
    """The library copies P * K values from each host array: the binding
    rejects arrays of another shape before any call (a short array would be
    read past its end)."""
    P, K = 5, 7
    ok = np.ones((P, K), np.float32)
    ctf, sig, dat = thx.particle_arrays(P, K, ok, ok, np.zeros((P, K, 2)))
    assert ctf.dtype == np.float32 and dat.shape == (P, K, 2)
    with pytest.raises(ValueError, match="ctf"):
        thx.particle_arrays(P, K, np.ones((P, K - 1)), ok)
    with pytest.raises(ValueError, match="sig"):
        thx.particle_arrays(P, K, ok, np.ones((P - 1, K)))
    with pytest.raises(ValueError, match="sig"):
        thx.particle_arrays(P, K, ok, np.ones(P * K))
    with pytest.raises(ValueError, match="dat"):
        thx.particle_arrays(P, K, ok, ok, np.zeros((P, K - 1, 2)))


def test_cpp_adapter_is_in_the_library():
    """The C++ adapter of include/thunder_expect.hpp is compiled into the
    library: every class it declares has its members exported."""
    out = subprocess.run(["nm", "-DC", "--defined-only", thx.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ["thx::Device::Device(int)", "thx::Device::setPixels(", "thx::Device::setParticles(",
                "thx::Projector::setProjectee(", "thx::Projector::setProjectee3D(",
                "thx::Projector::project(", "thx::Expectation::setGlobalGrid(",
                "thx::Expectation::expectation(", "thx::ManagedCalPoint::Init(",
                "thx::ExpectLocalRTD(", "thx::ExpectLocalPreI3D(", "thx::ExpectLocalM("]:
        assert sym in out, sym
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "adapter_check"))
