"""Projectee preparation on the GPU (needs gfx950): Projector::setProjectee
(src/Projector.cpp:123-148) with the LINEAR gridCorrection
(src/Projector.cpp:524-606), against the FP64 numpy oracle on the same map.

Inputs are the half spectra of seeded random real maps (so the half spectrum
is consistent and every correct inverse transform agrees).  The GPU runs both
FFTs and the correction in FP32: the padded spectrum must agree with the
oracle to 2e-5 of its largest magnitude, at power-of-two sizes, at sizes with
factors 3 and 5 (the box-160 and box-192 maps have them), and at box 128's
full size.  A second call on the same class with another size replaces the
volume.
"""
import numpy as np
import pytest

import oracle
import thx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pctx(gpu_ctx):
    """A context of its own: these volumes have other sizes than the pixel
    sets of the shared context."""
    ctx = thx.Context(0)
    yield ctx
    ctx.close()


def _map_ft(N, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(N, N, N))
    return np.fft.rfftn(f, axes=(0, 1, 2))


@pytest.mark.parametrize("N,pf", [(32, 2), (48, 2), (40, 2), (64, 1), (128, 2)])
def test_projectee_matches_oracle(pctx, N, pf):
    print(f"\n[projectee] N={N} pf={pf}", flush=True)
    ft = _map_ft(N, 100 + N)
    ms = pctx.set_projectee(0, ft, N, pf)
    V = pf * N
    got = pctx.get_volume(0, V)
    got = got[..., 0].astype(np.float64) + 1j * got[..., 1]
    ref = oracle.projectee(ft, N, pf)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"max error / max |F| = {err:.2e}; inverse FFT {ms[0]:.3f} ms, pad + correction "
          f"{ms[1]:.3f} ms, forward FFT {ms[2]:.3f} ms")
    assert err <= 2e-5


def test_projectee_replaces_class_volume(pctx):
    ft = _map_ft(16, 1)
    pctx.set_projectee(1, ft, 16, 2)
    ft2 = _map_ft(24, 2)
    pctx.set_projectee(1, ft2, 24, 2)
    got = pctx.get_volume(1, 48)
    got = got[..., 0].astype(np.float64) + 1j * got[..., 1]
    ref = oracle.projectee(ft2, 24, 2)
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 2e-5
