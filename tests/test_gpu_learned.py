"""Learned-proposal expectation (grid-restricted) on the GPU (needs gfx950).

* Whole selection (every grid rotation and translation selected): the gather
  scorer then searches the whole grid with BF16 operands, so the pose it
  returns must score, under the FP64 oracle, within 2e-3 relative of the
  oracle's best pose, and agree with that pose for >= 90% of particles (the
  bar of the BF16 global scan).
* Forced fallback (spread threshold below 1, which every proposal exceeds):
  each particle is rescored by the tiled whole-grid scan (BF16 or FP8) on
  gathered rows, whose results must equal those of the global scan bit for
  bit; with about half the particles flagged, their results must equal a
  fresh context's scan of just those particles, bit for bit.
* Selection and statistics from the network's mixture, against numpy FP64:
  the best rotation ranks within the nSelR most probable grid rotations; the
  tail depth (max log q - log q(best)) agrees to 1e-3 absolute and the spread
  (sum q)^2 / sum q^2 to 1e-3 relative (the kernel keeps log q in FP32).
* Gather-scorer shapes: whole selections of 77 .. 600 rotations and 9, 49
  and 81 translations run every rotation block (128 and 256 per block,
  ragged last blocks) and translation tile (32, 64, 128 rows); each result
  is held to the whole-selection bars against the FP64 oracle.
* BF16 proposal passes: the mixture from BF16 operands stays within stated
  bounds of the same network's FP32-operand mixture.
* Fallback by a wider selection: flagged particles get exactly the pose and
  score of the wider selection scored for every particle, the others those
  of the first selection (bit for bit), with half and with all flagged.
* Particle order of the selected-pose scorer (blocks dealt in proposal-mode
  order, or in index order): every result equal bit for bit, for one, two
  and three rotation blocks per particle.
* What training buys: at a fixed small selection the trained network's
  assignments agree with the plain FP32 scan on >= 80% of particles and at
  least 0.3 more often than an untrained network's.
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


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_forced_fallback_equals_global_scan(gpu_ctx, setup, prec):
    """Every particle flagged: the fallback is the whole-grid scan of that
    precision, bit for bit."""
    info, dvp, net = setup
    nf = gpu_ctx.expect_learned(net, 0, 32, 16, spread_max=0.5,
                                fallback_precision=thx.PRECISIONS[prec])
    assert nf == P
    bR, bT, sc, ess, spread, tail, fell = gpu_ctx.get_learned()
    assert fell.all()
    gpu_ctx.score_global(0, thx.PRECISIONS[prec])
    gR, gT, gS = gpu_ctx.get_best(0)
    np.testing.assert_array_equal(bR, gR)
    np.testing.assert_array_equal(bT, gT)
    np.testing.assert_array_equal(sc, gS)


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_partial_fallback_equals_subset_scan(gpu_ctx, setup, prec):
    """About half the particles flagged: their fallback results equal, bit for
    bit, a whole-grid scan of a fresh context holding only those particles.
    This holds the pieces the fallback reuses from the selection step (the
    BF16 slice table, the slices' per-shell maxima) and its split-K norm term
    to what a cold scan computes."""
    info, dvp, net = setup
    gpu_ctx.expect_learned(net, 0, 32, 16)
    spread = gpu_ctx.get_learned()[4]
    thr = float(np.median(spread))
    nf = gpu_ctx.expect_learned(net, 0, 32, 16, spread_max=thr,
                                fallback_precision=thx.PRECISIONS[prec])
    bR, bT, sc, ess, spread2, tail, fell = gpu_ctx.get_learned()
    idx = np.flatnonzero(fell)
    assert nf == len(idx) and 0 < nf < P
    dat, ctf, sig = gpu_ctx.get_particles()
    with thx.Context(0) as sub:
        sub.set_pixels(info["iCol"], info["iRow"], info["iSig"], BOX, PF)
        sub.set_volume(0, gpu_ctx.get_volume(0, V), V)
        sub.set_particles(dat[idx], ctf[idx], sig[idx])
        sub.set_grid(info["quat"], info["trans"])
        sub.score_global(0, thx.PRECISIONS[prec])
        gR, gT, gS = sub.get_best(0)
    np.testing.assert_array_equal(bR[idx], gR)
    np.testing.assert_array_equal(bT[idx], gT)
    np.testing.assert_array_equal(sc[idx], gS)


def test_bf16_proposal_matches_fp32(gpu_ctx, setup):
    """The proposal passes with BF16 operands (the default) against the same
    network with FP32 operands, on what the selection uses: the 32 most
    probable grid rotations shared on average for >= 90 % (>= 80 % for every
    particle; measured 99.1 % and 93.8 %), the log density at the FP32 pass's
    32 most probable rotations within 0.5 nat (measured 0.30: a component
    mean moves by up to a few degrees when its raw quaternion output is
    short), component weights within 0.01 and log concentrations within 0.05
    (measured 1.6e-3 and 8.9e-3).  FP32 accumulation throughout; BF16 keeps
    8 significant bits per operand."""
    info, dvp, net = setup
    C = net.C
    quat = info["quat"]
    net.set_precision(thx.FP32)
    m32 = net.propose(gpu_ctx)
    net.set_precision(thx.BF16)
    m16 = net.propose(gpu_ctx)
    dlq, overlap = [], []
    for p in range(P):
        a, b = _logq(m32[p], quat, C), _logq(m16[p], quat, C)
        top = np.argsort(-a)[:32]
        dlq.append(np.max(np.abs(a[top] - b[top])))
        overlap.append(len(np.intersect1d(top, np.argsort(-b)[:32])) / 32)
    q32 = m32[:, :6 * C].reshape(P, C, 6)
    q16 = m16[:, :6 * C].reshape(P, C, 6)
    dlk = np.abs(np.log(q32[..., 4]) - np.log(q16[..., 4]))
    dpi = np.abs(q32[..., 5] - q16[..., 5])
    print(f"max |d log q| at the top 32 {max(dlq):.3e}, top-32 overlap {np.mean(overlap):.3f} "
          f"(min {min(overlap):.3f}), log k {dlk.max():.2e}, pi {dpi.max():.2e}")
    assert max(dlq) <= 0.5
    assert np.mean(overlap) >= 0.9
    assert min(overlap) >= 0.8
    assert dpi.max() <= 0.01
    assert dlk.max() <= 0.05


def test_whole_selection_large_grid():
    """A 9,000-rotation grid: the selection kernel's per-particle log
    densities need 36 KB of dynamic LDS (beyond the default limit, raised at
    launch).  With the whole grid selected, the gather scorer's best poses
    agree with the BF16 global scan's for >= 90 % of particles and its best
    scores to 2e-3 relative."""
    R9 = 9000
    print(f"\n[case] box64 P={P} R={R9} whole selection", flush=True)
    with thx.Context(0) as ctx:
        info = wl.setup_context(ctx, "box64", n_particles=P, n_rot=R9)
        T = len(info["trans"])
        net = thx.Proposal(components=4, seed=3)
        try:
            nf = ctx.expect_learned(net, 0, R9, T)
        finally:
            net.close()
        assert nf == 0
        bR, bT, sc = ctx.get_learned()[:3]
        ctx.score_global(0, thx.BF16)
        gR, gT, gS = ctx.get_best(0)
    agree = np.mean((bR == gR) & (bT == gT))
    rel = np.max(np.abs(sc - gS) / np.abs(gS))
    print(f"agreement {agree:.3f}, max score difference {rel:.2e}")
    assert agree >= 0.9
    assert rel <= 2e-3


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


def test_wide_fallback_equals_wide_selection():
    print("\n[setup] box64 P=60 R=600 T=49 wide fallback", flush=True)
    ctx = thx.Context(0)
    wl.setup_context(ctx, "box64", n_particles=60, n_rot=600)
    net = thx.Proposal(components=4, seed=3)
    try:
        ctx.set_learned_fallback_selection(0, 0)
        ctx.expect_learned(net, 0, 300, 49)
        wide = ctx.get_learned()
        ctx.expect_learned(net, 0, 64, 25)
        first = ctx.get_learned()
        spread = first[4]
        ctx.set_learned_fallback_selection(300, 49)
        for smax in (float(np.median(spread)), 0.5):
            nf = ctx.expect_learned(net, 0, 64, 25, 0.0, smax, 0.0, 0.0, thx.BF16)
            got = ctx.get_learned()
            fell = got[6].astype(bool)
            assert nf == int(fell.sum()) and nf > 0
            assert np.array_equal(fell, ~(spread <= smax))
            for i in (0, 1, 2):   # best rotation, translation, score
                assert np.array_equal(got[i][fell], wide[i][fell]), (smax, i)
                assert np.array_equal(got[i][~fell], first[i][~fell]), (smax, i)
    finally:
        ctx.set_learned_fallback_selection(0, 0)
        net.close()
        ctx.close()


def test_select_order_does_not_change_results():
    print("\n[setup] box64 P=60 R=600 T=49 select order", flush=True)
    ctx = thx.Context(0)
    wl.setup_context(ctx, "box64", n_particles=60, n_rot=600)
    net = thx.Proposal(components=4, seed=3)
    try:
        for nr, nt in ((600, 49), (300, 25), (160, 9), (100, 25)):
            out = {}
            for on in (True, False):
                ctx.set_select_order(on)
                nf = ctx.expect_learned(net, 0, nr, nt, 0.0, 0.0, 0.0, 0.0, thx.BF16)
                out[on] = (nf,) + tuple(ctx.get_learned())
            for a, b in zip(out[True], out[False]):
                assert np.array_equal(a, b), (nr, nt)
    finally:
        ctx.set_select_order(True)
        net.close()
        ctx.close()


def test_trained_network_gains(gpu_ctx):
    """What training buys: with the same small selection (10% of the rotations,
    half the translations), the trained network's assignments agree with the
    plain FP32 whole-grid scan far more often than an untrained network's."""
    import learned
    print("\n[setup] box64 P=300 R=1000 T=49 training", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=300, n_rot=1000)
    quat = info["quat"]
    step = wl.sampling_step(quat)
    gpu_ctx.score_global(0, thx.FP32)
    gR, gT, _ = gpu_ctx.get_best(0)
    w = wl.WORKLOADS["box64"]
    tctx = thx.Context(0)
    tctx.set_pixels(info["iCol"], info["iRow"], info["iSig"], w["box"], w["pf"])
    tctx.synth_volume(0, w["box"], w["pf"], w["n_blob"], w["blob_sigma"], w["blob_radius"],
                      wl.seeds("box64")["volume"])
    agree = {}
    for tag, n_samples in (("untrained", 0), ("trained", 40 * 8192)):
        net = thx.Proposal(components=8, seed=11)
        if n_samples:
            secs, _ = learned.train_proposal(net, tctx, "box64", n_samples=n_samples)
            print(f"trained on {n_samples} samples in {secs:.1f} s")
        gpu_ctx.expect_learned(net, 0, 100, 25)
        bR = gpu_ctx.get_learned()[0]
        agree[tag] = float(np.mean(wl.geodesic_angle(quat[bR], quat[gR]) <= step * (1 + 1e-6)))
        net.close()
    tctx.close()
    print(f"agreement within one sampling step: {agree}")
    assert agree["trained"] >= 0.8
    assert agree["trained"] - agree["untrained"] >= 0.3


@pytest.mark.parametrize("R,side", [(77, 3), (200, 7), (600, 7), (600, 3), (300, 9)])
def test_whole_selection_shapes(gpu_ctx, R, side):
    Pn = 16
    print(f"\n[setup] box64 P={Pn} R={R} T={side * side} learned whole selection", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=Pn, n_rot=R)
    trans = wl.shift_grid(side, 1.0)
    gpu_ctx.set_grid(info["quat"], trans)
    T = len(trans)
    vol = gpu_ctx.get_volume(0, V)
    dat, ctf, sig = gpu_ctx.get_particles()
    dvp = oracle.global_scan(vol, V, PF, info["iCol"], info["iRow"], BOX, info["quat"], trans,
                             dat, ctf, sig, np.full(R, 1.0 / R), np.full(T, 1.0 / T), f64=True,
                             want_dvp=True)[4]
    net = thx.Proposal(components=4, seed=3)
    try:
        nf = gpu_ctx.expect_learned(net, 0, R, T)
        bR, bT, sc, *_ = gpu_ctx.get_learned()
    finally:
        net.close()
    assert nf == 0
    flat = dvp.reshape(Pn, -1)
    best = flat.max(1)
    chosen = flat[np.arange(Pn), bR * T + bT]
    agree = np.mean(flat.argmax(1) == bR * T + bT)
    rel_gap = (best - chosen) / np.abs(best)
    rel_score = np.abs(sc - best) / np.abs(best)
    print(f"agreement {agree:.3f}, max FP64 gap {rel_gap.max():.2e}, "
          f"max score err {rel_score.max():.2e}")
    assert agree >= 0.85
    assert rel_gap.max() <= 2e-3
    assert rel_score.max() <= 2e-3


def test_seeded_forced_fallback_equals_global_path(gpu_ctx, setup):
    """Proposal-seeded local search with the ESS threshold above any ESS:
    every particle takes the uniform global search, so the result must equal,
    bit for bit, the reference's own path on the same particles: whole-grid
    scan, particle states from it (thx_pf_from_global with the fallback's
    seed), local phases."""
    info, dvp, net = setup
    # the module's particles and grid again (other tests here replace them)
    wl.setup_context(gpu_ctx, "box64", n_particles=P, n_rot=R)
    args = dict(min_std_r=192 ** (-1.0 / 3), min_std_t=0.3, pf_l=2.0, pf_s=0.5, transS=1.0,
                max_phase=12)
    keys = ("quat", "trans", "wR", "wT", "k", "s01", "topR", "topT", "peakR", "nPhase")
    nf, ph = gpu_ctx.expect_seeded(net, 0, 100, 25, thx.BF16, ess_min=1e9, fallback=thx.BF16,
                                   seed=5, **args)
    assert nf == P and ph >= 10
    st = gpu_ctx.pf_get_state()
    gpu_ctx.score_global(0, thx.BF16)
    gpu_ctx.pf_from_global(0, 100, 25, args["min_std_r"], args["min_std_t"], args["pf_s"], 1.0,
                           seed=5 ^ 0x5eed)
    ph2 = gpu_ctx.pf_refine(None, thx.BF16, after_global=True, pf_l=2.0, pf_s=0.5, max_phase=12,
                            seed=6)
    ref = gpu_ctx.pf_get_state()
    assert ph2 == ph
    for k in keys:
        np.testing.assert_array_equal(st[k], ref[k], err_msg=k)


def test_seeded_local_search_refines():
    """Proposal-seeded local search without fallback, against the reference's
    flow on the same particles (whole-grid FP32 scan, states from it, local
    phases): with a trained network the seeded search converges and its final
    poses are as close to the truth, far below the grid's spacing, without the
    whole-grid scan."""
    import learned
    print("\n[setup] box64 P=128 R=1000 T=49 seeded local search", flush=True)
    w = wl.WORKLOADS["box64"]
    with thx.Context(0) as ctx, thx.Context(0) as tctx:
        info = wl.setup_context(ctx, "box64", n_particles=128, n_rot=1000)
        step = wl.sampling_step(info["quat"])
        tctx.set_pixels(info["iCol"], info["iRow"], info["iSig"], w["box"], w["pf"])
        tctx.synth_volume(0, w["box"], w["pf"], w["n_blob"], w["blob_sigma"], w["blob_radius"],
                          wl.seeds("box64")["volume"])
        net = thx.Proposal(components=8, seed=11)
        learned.train_proposal(net, tctx, "box64", n_samples=60 * 8192)
        args = dict(min_std_r=1000 ** (-1.0 / 3), min_std_t=0.3, pf_l=2.0, pf_s=0.5,
                    transS=1.0, max_phase=30)
        nf, ph = ctx.expect_seeded(net, 0, 300, 25, thx.BF16, ess_min=0.0, seed=21, **args)
        st = ctx.pf_get_state()
        net.close()
        ctx.score_global(0, thx.FP32)
        ctx.pf_from_global(0, 300, 25, args["min_std_r"], args["min_std_t"], args["pf_s"], 1.0,
                           seed=21)
        ctx.pf_refine(None, thx.BF16, after_global=True, pf_l=2.0, pf_s=0.5, max_phase=30,
                      seed=22)
        ref = ctx.pf_get_state()
    err = wl.geodesic_angle(st["topR"], info["true_quat"])
    err_ref = wl.geodesic_angle(ref["topR"], info["true_quat"])
    print(f"seeded: {ph} phases, converged {np.mean(st['nPhase'] >= 0):.2f}, median error "
          f"{np.degrees(np.median(err)):.2f} deg; global scan + local: "
          f"{np.degrees(np.median(err_ref)):.2f} deg (grid step {np.degrees(step):.1f})")
    assert nf == 0
    assert np.mean(st["nPhase"] >= 0) >= 0.9
    assert np.median(err) < 0.25 * step
    assert np.median(err) <= 1.5 * np.median(err_ref) + np.radians(0.2)
