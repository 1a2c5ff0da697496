"""The entry point that allocates nothing (thx_pf_resample_noalloc), called
through the C ABI on device buffers, against thx_pf_resample on the same
inputs: indices, weights and ESS must be identical (integer indices exactly,
the FP64 weights bit for bit, since both run the same kernel on the same
values).  Two calls in a row, the second with fewer particles, so the second
call runs over the offsets the first one left in the shared scratch array.

Device memory for the inputs and outputs comes from the HIP runtime the
library itself is bound to, found by its soname among the loaded libraries,
through ctypes.
"""
import ctypes as C

import numpy as np
import pytest

import thx

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    def __init__(self):
        thx.load()   # loads the HIP runtime as the library's dependency
        # by soname: the runtime already loaded for the library, not a second copy
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D) == 0
        return p

    def empty(self, n, dtype):
        return self.put(np.full(n, -7, dtype))

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def _batch(rng, P):
    nIn = rng.integers(1, 3000, P).astype(np.int32)
    nIn[0] = 8192
    nOut = rng.integers(1, 2000, P).astype(np.int32)
    nOut[-1] = 8192
    w = rng.uniform(0.1, 1.0, int(nIn.sum()))
    u = np.exp(rng.normal(0, 3, int(nIn.sum())))
    return nIn, nOut, w, u


def _noalloc(hip, ctx, w, u, nIn, nOut, seed, u0):
    P, tout = nIn.size, int(nOut.sum())
    dw, du, dni, dno = hip.put(w), hip.put(u), hip.put(nIn), hip.put(nOut)
    du0 = None if u0 is None else hip.put(u0)
    didx, dwn, dess = hip.empty(tout, np.int32), hip.empty(tout, np.float64), hip.empty(P, np.float64)
    st = thx.load().thx_pf_resample_noalloc(ctx._h, dw, du, dni, dno, P, seed, du0, didx, dwn, dess)
    thx.check(st, "pf_resample_noalloc")
    ctx.sync()
    return (hip.get(didx, tout, np.int32), hip.get(dwn, tout, np.float64),
            hip.get(dess, P, np.float64))


def test_noalloc_resample_matches_resample(gpu_ctx):
    hip = _Hip()
    try:
        rng = np.random.default_rng(21)
        # call 1: 300 particles with given offsets u0
        nIn, nOut, w, u = _batch(rng, 300)
        u0 = rng.uniform(0, 1, nIn.size) / nOut
        print(f"\n[noalloc] call 1: P={nIn.size}, {nIn.sum()} samples in, {nOut.sum()} out",
              flush=True)
        got = _noalloc(hip, gpu_ctx, w, u, nIn, nOut, 5, u0)
        ref = gpu_ctx.pf_resample(w, u, nIn, nOut, seed=5, u0=u0)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        # call 2: fewer particles, offsets drawn from the seed on the device
        nIn, nOut, w, u = _batch(rng, 40)
        print(f"[noalloc] call 2: P={nIn.size}, {nIn.sum()} samples in, {nOut.sum()} out",
              flush=True)
        got = _noalloc(hip, gpu_ctx, w, u, nIn, nOut, 9, None)
        ref = gpu_ctx.pf_resample(w, u, nIn, nOut, seed=9)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        # a particle with a count outside [1, 8192] is skipped (ess = -1);
        # the others are as without it
        nIn, nOut, w, u = _batch(rng, 6)
        u0 = rng.uniform(0, 1, 6) / nOut
        bad = nIn.copy()
        bad[2] = 9000
        w2 = np.concatenate([w[:nIn[:2].sum()], np.ones(9000), w[nIn[:3].sum():]])
        u2 = np.concatenate([u[:nIn[:2].sum()], np.ones(9000), u[nIn[:3].sum():]])
        # the skipped particle's input slots are counted as 8192
        keep = np.r_[0:nIn[:2].sum(), nIn[:2].sum():nIn[:2].sum() + 8192,
                     nIn[:2].sum() + 9000:w2.size]
        gi, gw, ge = _noalloc(hip, gpu_ctx, w2[keep], u2[keep], bad, nOut, 3, u0)
        assert ge[2] == -1.0
        sel = np.r_[0:2, 3:6]
        ri, rw, re_ = gpu_ctx.pf_resample(w[np.r_[0:nIn[:2].sum(), nIn[:3].sum():w.size]],
                                          u[np.r_[0:nIn[:2].sum(), nIn[:3].sum():u.size]],
                                          nIn[sel], nOut[sel], u0=u0[sel])
        o = np.r_[0, np.cumsum(nOut)]
        mask = np.ones(o[-1], bool)
        mask[o[2]:o[3]] = False
        np.testing.assert_array_equal(gi[mask], ri)
        np.testing.assert_array_equal(gw[mask], rw)
        np.testing.assert_array_equal(ge[sel], re_)
        assert np.all(gi[~mask] == -7)            # its outputs are left untouched
    finally:
        hip.free()


def test_noalloc_rejects_more_particles_than_the_scratch(gpu_ctx):
    lib = thx.load()
    P = 131072 + 1
    st = lib.thx_pf_resample_noalloc(gpu_ctx._h, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                                     C.c_void_p(8), P, 1, None, C.c_void_p(8), C.c_void_p(8), None)
    assert st == 5 and b"scratch" in lib.thx_last_error()
