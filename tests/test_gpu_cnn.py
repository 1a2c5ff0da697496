"""Learned-proposal CNN kernels against the FP64 oracle (needs gfx950).

Tolerances: images to 5e-4 of their unit scale (FP32 sums of ~500 terms), the
BF16 transform of the proposal passes to 0.01 RMS and 0.05 max of it (three
BF16 roundings of 2^-9 relative: coefficients, twiddles, the first pass's
output); forward outputs to 1e-4
relative; loss to 1e-5 relative; parameter gradients, per tensor, to 2e-3 of
that tensor's largest gradient (FP32 accumulation over up to 2.6e5 terms).
"""
import numpy as np
import pytest

import cnn_oracle as co
import thx
import workloads as wl

pytestmark = pytest.mark.gpu
C = 8


def test_images_match_numpy(gpu_ctx):
    print("\n[setup] box64 P=6 images", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=6, n_rot=16)
    dat, ctf, _ = gpu_ctx.get_particles()
    img = thx.ctx_images(gpu_ctx)
    iCol, iRow = info["iCol"], info["iRow"]
    sel = (iCol < 32) & (iRow >= -32) & (iRow < 32) & ~((iCol == 0) & (iRow == 0))
    y, x = np.mgrid[0:64, 0:64]
    for p in range(6):
        X = np.where(ctf[p, sel] >= 0, 1, -1) * (dat[p, sel, 0] + 1j * dat[p, sel, 1])
        ph = 2j * np.pi * (iCol[sel][:, None, None] * x + iRow[sel][:, None, None] * y) / 64
        im = 2 * np.real(np.sum(X[:, None, None] * np.exp(ph), 0))
        im = (im - im.mean()) / im.std()
        assert np.abs(img[p] - im).max() < 5e-4


def test_images_bf16_match_numpy(gpu_ctx):
    print("\n[setup] box64 P=6 images, BF16 transform", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=6, n_rot=16)
    dat, ctf, _ = gpu_ctx.get_particles()
    img = thx.ctx_images(gpu_ctx, thx.BF16)
    iCol, iRow = info["iCol"], info["iRow"]
    sel = (iCol < 32) & (iRow >= -32) & (iRow < 32) & ~((iCol == 0) & (iRow == 0))
    y, x = np.mgrid[0:64, 0:64]
    for p in range(6):
        X = np.where(ctf[p, sel] >= 0, 1, -1) * (dat[p, sel, 0] + 1j * dat[p, sel, 1])
        ph = 2j * np.pi * (iCol[sel][:, None, None] * x + iRow[sel][:, None, None] * y) / 64
        im = 2 * np.real(np.sum(X[:, None, None] * np.exp(ph), 0))
        im = (im - im.mean()) / im.std()
        d = img[p] - im
        print(f"particle {p}: rms {np.sqrt(np.mean(d * d)):.2e} max {np.abs(d).max():.2e}")
        assert np.sqrt(np.mean(d * d)) < 0.01
        assert np.abs(d).max() < 0.05


@pytest.fixture(scope="module")
def net():
    n = thx.Proposal(components=C, seed=42)
    yield n
    n.close()


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, 64, 64)).astype(np.float32)
    q = rng.normal(size=(B, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(0, 2, size=(B, 2))
    return x, q, t


def test_forward_and_gradient_match_oracle(net):
    x, q, t = _batch(4, 0)
    p = net.get_params().astype(np.float64)
    out = net.forward(x)
    ref_loss, ref_g, ref_out = co.loss_grad(p, x[:, None].astype(np.float64), q, t, C)
    np.testing.assert_allclose(out, ref_out, rtol=1e-4, atol=1e-4 * np.abs(ref_out).max())
    loss, g = net.loss_grad(x, q, t)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in co.layout(C)])
    worst = 0.0
    for a, b in zip(offs[:-1], offs[1:]):
        scale = np.abs(ref_g[a:b]).max()
        err = np.abs(g[a:b] - ref_g[a:b]).max() / scale
        worst = max(worst, err)
        assert err < 2e-3, (a, b, err)
    print(f"max per-tensor relative gradient error {worst:.2e}")


def test_training_reduces_loss(net):
    x, q, t = _batch(64, 1)
    # learn a fixed small set: loss must fall well below its start
    l0 = net.train_step(x, q, t, 1e-3)
    for _ in range(60):
        l = net.train_step(x, q, t, 1e-3)
    print(f"loss {l0:.3f} -> {l:.3f}")
    assert l < l0 - 1.0


def test_training_is_deterministic():
    """Same seed, same batches: bit-identical parameters (no atomics in the
    gradient reductions)."""
    x, q, t = _batch(64, 2)
    params = []
    for _ in range(2):
        n = thx.Proposal(components=C, seed=5)
        for _ in range(3):
            n.train_step(x, q, t, 1e-3)
        params.append(n.get_params())
        n.close()
    np.testing.assert_array_equal(params[0], params[1])
