"""Learned-proposal expectation (grid-restricted) on the GPU (needs gfx950).

* Whole selection (every grid rotation and translation selected): the gather
  scorer then searches the whole grid with BF16 operands, so the pose it
  returns must score, under the FP64 oracle, within 2e-3 relative of the
  oracle's best pose, and agree with that pose for >= 90% of particles (the
  bar of the BF16 global scan).
* Forced fallback (spread threshold below 1, which every proposal exceeds):
  each particle is rescored by the tiled whole-grid BF16 scan on gathered
  rows, whose results must equal those of the global BF16 scan bit for bit.
* Selection and statistics from the network's mixture, against numpy FP64:
  the best rotation ranks within the nSelR most probable grid rotations; the
  tail depth (max log q - log q(best)) agrees to 1e-3 absolute and the spread
  (sum q)^2 / sum q^2 to 1e-3 relative (the kernel keeps log q in FP32).
"""
import numpy as np
import pytest

import oracle
import thx
import workloads as wl

pytestmark = pytest.mark.gpu

BOX, PF, V = 64, 2, 128
P, R = 40, 192


@pytest.fixture(scope="module")
def setup(gpu_ctx):
    print("\n[setup] box64 P=40 R=192 T=49 learned", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=P, n_rot=R)
    vol = gpu_ctx.get_volume(0, V)
    dat, ctf, sig = gpu_ctx.get_particles()
    T = len(info["trans"])
    dvp = oracle.global_scan(vol, V, PF, info["iCol"], info["iRow"], BOX, info["quat"],
                             info["trans"], dat, ctf, sig, np.full(R, 1.0 / R),
                             np.full(T, 1.0 / T), f64=True, want_dvp=True)[4]
    net = thx.Proposal(components=4, seed=3)
    yield info, dvp, net
    net.close()


def _logq(mix, quat, C):
    lp = []
    for c in range(C):
        m, k, pi = mix[6 * c:6 * c + 4], mix[6 * c + 4], mix[6 * c + 5]
        d = quat @ m
        s = d * d + (1 - d * d) / k
        lp.append(np.log(pi) - np.log(2 * np.pi ** 2) - 1.5 * np.log(k) - 2 * np.log(s))
    lp = np.array(lp)
    mx = lp.max(0)
    return mx + np.log(np.exp(lp - mx).sum(0))


def test_whole_selection_matches_oracle(gpu_ctx, setup):
    info, dvp, net = setup
    T = len(info["trans"])
    nf = gpu_ctx.expect_learned(net, 0, R, T)
    assert nf == 0
    bR, bT, sc, ess, spread, tail, fell = gpu_ctx.get_learned()
    flat = dvp.reshape(P, -1)
    best = flat.max(1)
    chosen = flat[np.arange(P), bR * T + bT]
    agree = np.mean(flat.argmax(1) == bR * T + bT)
    rel_gap = (best - chosen) / np.abs(best)
    rel_score = np.abs(sc - best) / np.abs(best)
    print(f"agreement {agree:.3f}, max FP64 gap {rel_gap.max():.2e}, "
          f"max score err {rel_score.max():.2e}")
    assert agree >= 0.9
    assert rel_gap.max() <= 2e-3
    assert rel_score.max() <= 2e-3
    assert not fell.any()


def test_forced_fallback_equals_global_bf16(gpu_ctx, setup):
    info, dvp, net = setup
    nf = gpu_ctx.expect_learned(net, 0, 32, 16, spread_max=0.5, fallback_precision=thx.BF16)
    assert nf == P
    bR, bT, sc, ess, spread, tail, fell = gpu_ctx.get_learned()
    assert fell.all()
    gpu_ctx.score_global(0, thx.BF16)
    gR, gT, gS = gpu_ctx.get_best(0)
    np.testing.assert_array_equal(bR, gR)
    np.testing.assert_array_equal(bT, gT)
    np.testing.assert_array_equal(sc, gS)


def test_selection_and_statistics(gpu_ctx, setup):
    info, dvp, net = setup
    nsel = 24
    gpu_ctx.expect_learned(net, 0, nsel, 9)
    bR, bT, sc, ess, spread, tail, fell = gpu_ctx.get_learned()
    mix = net.propose(gpu_ctx)
    C = net.C
    for p in range(P):
        lq = _logq(mix[p], info["quat"], C)
        rank = np.sum(lq > lq[bR[p]])
        assert rank < nsel + 2, (p, rank)
        assert abs(tail[p] - (lq.max() - lq[bR[p]])) <= 1e-3, p
        q = np.exp(lq - lq.max())
        ref_spread = q.sum() ** 2 / np.sum(q * q)
        assert abs(spread[p] - ref_spread) <= 1e-3 * ref_spread, p
        # translation: within the 9 nearest grid shifts to the predicted mean
        d2 = np.sum((info["trans"] - mix[p, 6 * C:6 * C + 2]) ** 2, 1)
        assert np.sum(d2 < d2[bT[p]] - 1e-9) < 9, p
    assert np.all((ess >= 1.0 - 1e-9) & (ess <= nsel + 1e-9))
