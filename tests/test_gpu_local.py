"""Ragged local-search scorer against the FP64 oracle (needs gfx950).

Tolerances: FP32 baseline (best log-likelihood) relative error <= 1e-4 and the
same best (rotation, translation) per particle; marginals within 3e-3 of each
particle's largest entry (an FP32 log-likelihood of magnitude ~1e3 carries an
absolute rounding error ~1e-3, which is the relative error of exp(dvp - max)).
BF16 and FP16 (FP16 with per-shell operand scales taken from the class
volumes and the particles): best-score relative error <= 2e-3 and >= 80%
identical best poses on this small noisy case.
"""
import numpy as np
import pytest

import oracle
import thx
import workloads as wl

pytestmark = pytest.mark.gpu

BOX, PF = 64, 2
V = BOX * PF


def _batch(info, P, rng, rmax, tmax):
    nR = rng.integers(20, rmax + 1, P).astype(np.int32)
    nR[0] = rmax                      # at least one particle spans several blocks
    nT = rng.integers(1, tmax + 1, P).astype(np.int32)
    nT[1] = tmax
    quats, trans = [], []
    for p in range(P):
        q0 = info["true_quat"][p]
        # half near the true pose, half anywhere
        qn = q0 + 0.05 * rng.normal(size=(nR[p], 4))
        qa = rng.normal(size=(nR[p], 4))
        q = np.where((np.arange(nR[p]) % 2 == 0)[:, None], qn, qa)
        quats.append(q / np.linalg.norm(q, axis=1, keepdims=True))
        trans.append(info["true_shift"][p] + rng.uniform(-2, 2, size=(nT[p], 2)))
    pR = rng.uniform(0.5, 1.5, int(nR.sum()))
    pT = rng.uniform(0.5, 1.5, int(nT.sum()))
    return np.concatenate(quats), nR, np.concatenate(trans), nT, pR, pT


@pytest.fixture(scope="module")
def local_case(gpu_ctx):
    print("\n[setup] local box64 P=24", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=24, n_rot=64)
    vol = gpu_ctx.get_volume(0, V)
    dat, ctf, sig = gpu_ctx.get_particles()
    info.update(vol=vol, dat=dat, ctf=ctf, sig=sig)
    return info


@pytest.mark.parametrize("tmax", [25, 60, 100])
def test_local_fp32_matches_oracle(gpu_ctx, local_case, tmax):
    c = local_case
    rng = np.random.default_rng(tmax)
    P = 24
    quat, nR, trans, nT, pR, pT = _batch(c, P, rng, 300, tmax)
    print(f"[case] tmax={tmax} Rtot={nR.sum()} Ttot={nT.sum()}", flush=True)
    gpu_ctx.score_local(quat, nR, trans, nT, pR, pT, precision=thx.FP32)
    wC, wR, wT, base, bR, bT = gpu_ctx.get_local()
    ref = oracle.local_scan(c["vol"], V, PF, c["iCol"], c["iRow"], BOX, quat, nR, trans, nT,
                            pR, pT, c["dat"], c["ctf"], c["sig"])
    rel = np.abs(base - ref[3]) / np.abs(ref[3])
    print(f"FP32 local: max rel base err {rel.max():.2e}")
    assert rel.max() <= 1e-4
    np.testing.assert_array_equal(bR, ref[4])
    np.testing.assert_array_equal(bT, ref[5])
    np.testing.assert_allclose(wC, ref[0], rtol=3e-3)
    ro = np.r_[0, np.cumsum(nR)]
    to = np.r_[0, np.cumsum(nT)]
    for p in range(P):
        a, b = wR[ro[p]:ro[p + 1]], ref[1][ro[p]:ro[p + 1]]
        assert np.abs(a - b).max() <= 3e-3 * b.max()
        a, b = wT[to[p]:to[p + 1]], ref[2][to[p]:to[p + 1]]
        assert np.abs(a - b).max() <= 3e-3 * b.max()


@pytest.mark.parametrize("prec", [thx.BF16, thx.FP16])
def test_local_16bit_close_to_oracle(gpu_ctx, local_case, prec):
    c = local_case
    rng = np.random.default_rng(5)
    quat, nR, trans, nT, pR, pT = _batch(c, 24, rng, 260, 49)
    gpu_ctx.score_local(quat, nR, trans, nT, pR, pT, precision=prec)
    wC, wR, wT, base, bR, bT = gpu_ctx.get_local()
    ref = oracle.local_scan(c["vol"], V, PF, c["iCol"], c["iRow"], BOX, quat, nR, trans, nT,
                            pR, pT, c["dat"], c["ctf"], c["sig"])
    rel = np.abs(base - ref[3]) / np.abs(ref[3])
    agree = np.mean((bR == ref[4]) & (bT == ref[5]))
    print(f"precision {prec} local: max rel base err {rel.max():.2e}, best-pose agreement "
          f"{agree:.3f}")
    assert rel.max() <= 2e-3
    assert agree >= 0.8


def test_local_matches_global_on_shared_grid(gpu_ctx, local_case):
    """Local scan of every particle over the global grid equals the global scan."""
    c = local_case
    P = 24
    R, T = len(c["quat"]), len(c["trans"])
    quat = np.tile(c["quat"], (P, 1))
    trans = np.tile(c["trans"], (P, 1))
    nR = np.full(P, R, np.int32)
    nT = np.full(P, T, np.int32)
    gpu_ctx.score_local(quat, nR, trans, nT, precision=thx.FP32)
    wC, wR, wT, base, bR, bT = gpu_ctx.get_local()
    gpu_ctx.score_global(0, thx.FP32)
    gC, gR, gT, gb = gpu_ctx.get_global(1)
    gbR, gbT, _ = gpu_ctx.get_best(0)
    np.testing.assert_allclose(base, gb, rtol=1e-5)
    np.testing.assert_array_equal(bR, gbR)
    np.testing.assert_array_equal(bT, gbT)
    np.testing.assert_allclose(wC, gC[:, 0], rtol=1e-3)


def test_local_box160_three_classes(gpu_ctx):
    """The ragged three-class batch of the box-160 local search (cutoff 64,
    K = 6,323) at the workload's per-particle counts: 200-2,000 rotations drawn
    about the true pose (spread of a von Mises-Fisher kappa = 50) and 25-81
    translations, two particles per class, each scored against its own class
    volume in one call (one launch per translation tile).  FP32 against the FP64
    oracle run per class; BF16 and FP16 (the workload's cross-term format):
    best-score relative error <= 2e-3, the same best pose as the FP64 oracle for
    >= 5 of the 6 particles, and every chosen pose scoring, under the FP64
    oracle, within 2e-3 relative of that particle's FP64 optimum."""
    name = "box160"
    w = wl.WORKLOADS[name]
    sd = wl.seeds(name)
    iCol, iRow, iSig = wl.pixel_set(w["box"], w["cutoff"])
    assert len(iCol) == 6323
    gpu_ctx.set_pixels(iCol, iRow, iSig, w["box"], w["pf"])
    Vb = w["box"] * w["pf"]
    per, nCls = 2, 3
    parts, truth, vols = [], [], []
    for c in range(nCls):
        gpu_ctx.synth_volume(c, w["box"], w["pf"], w["n_blob"], w["blob_sigma"],
                             w["blob_radius"], sd["volume"] + c)
        q, t, _ = gpu_ctx.synth_particles(c, per, w["max_shift"], w["df_min"], w["df_max"],
                                          w["snr"], sd["particles"] + c, **wl.CTF_DEFAULTS)
        parts.append(gpu_ctx.get_particles())
        truth.append((q, t))
        vols.append(gpu_ctx.get_volume(c, Vb))
    dat = np.concatenate([x[0] for x in parts])
    ctf = np.concatenate([x[1] for x in parts])
    sig = np.concatenate([x[2] for x in parts])
    gpu_ctx.set_particles(dat, ctf, sig)
    P = per * nCls
    cls = np.repeat(np.arange(nCls), per).astype(np.int32)
    tq = np.concatenate([x[0] for x in truth])
    ts = np.concatenate([x[1] for x in truth])
    rng = np.random.default_rng(160)
    nR = rng.integers(200, 2001, P).astype(np.int32)
    nT = rng.integers(25, 82, P).astype(np.int32)
    nR[0], nT[1] = 2000, 81   # the largest counts of the workload
    quats, trans = [], []
    for p in range(P):
        q = tq[p] + rng.normal(0, 1 / np.sqrt(2 * 50), size=(nR[p], 4))
        quats.append(q / np.linalg.norm(q, axis=1, keepdims=True))
        trans.append(ts[p] + rng.uniform(-3, 3, size=(nT[p], 2)))
    quat, tr = np.concatenate(quats), np.concatenate(trans)
    pR = rng.uniform(0.5, 1.5, int(nR.sum()))
    pT = rng.uniform(0.5, 1.5, int(nT.sum()))
    print(f"\n[case] box160 P={P} Rtot={nR.sum()} Ttot={nT.sum()}", flush=True)
    ro = np.r_[0, np.cumsum(nR)]
    to = np.r_[0, np.cumsum(nT)]
    ref = [None] * P
    for c in range(nCls):
        ps = np.where(cls == c)[0]
        sel_r = np.concatenate([np.arange(ro[p], ro[p + 1]) for p in ps])
        sel_t = np.concatenate([np.arange(to[p], to[p + 1]) for p in ps])
        out = oracle.local_scan(vols[c], Vb, w["pf"], iCol, iRow, w["box"], quat[sel_r],
                                nR[ps], tr[sel_t], nT[ps], pR[sel_r], pT[sel_t], dat[ps],
                                ctf[ps], sig[ps])
        r0 = np.r_[0, np.cumsum(nR[ps])]
        t0 = np.r_[0, np.cumsum(nT[ps])]
        for i, p in enumerate(ps):
            ref[p] = (out[0][i], out[1][r0[i]:r0[i + 1]], out[2][t0[i]:t0[i + 1]], out[3][i],
                      out[4][i], out[5][i])
    for prec in (thx.FP32, thx.BF16, thx.FP16):
        gpu_ctx.score_local(quat, nR, tr, nT, pR, pT, cls=cls, precision=prec)
        wC, wR, wT, base, bR, bT = gpu_ctx.get_local()
        rel = np.array([abs(base[p] - ref[p][3]) / abs(ref[p][3]) for p in range(P)])
        same = np.mean([(bR[p], bT[p]) == (ref[p][4], ref[p][5]) for p in range(P)])
        print(f"precision {prec}: max rel base err {rel.max():.2e}, same best pose {same:.2f}")
        if prec == thx.FP32:
            assert rel.max() <= 1e-4 and same == 1.0
            for p in range(P):
                assert np.abs(wR[ro[p]:ro[p + 1]] - ref[p][1]).max() <= 3e-3 * ref[p][1].max()
                assert np.abs(wT[to[p]:to[p + 1]] - ref[p][2]).max() <= 3e-3 * ref[p][2].max()
        else:
            assert rel.max() <= 2e-3
            assert same >= 5 / 6
            # the FP64 score of each chosen pose against the FP64 optimum
            for p in range(P):
                c = cls[p]
                one = oracle.local_scan(vols[c], Vb, w["pf"], iCol, iRow, w["box"],
                                        quat[ro[p] + bR[p]][None], np.array([1], np.int32),
                                        tr[to[p] + bT[p]][None], np.array([1], np.int32),
                                        np.ones(1), np.ones(1), dat[p:p + 1], ctf[p:p + 1],
                                        sig[p:p + 1])
                gap = abs(one[3][0] - ref[p][3]) / abs(ref[p][3])
                assert gap <= 2e-3, (prec, p, gap)


def test_local_launch_shapes_and_graph_replays(gpu_ctx, local_case):
    """Direct and persistent launches give bit-identical results (the same
    per-item code); the autotuned shape is one of them; a captured graph,
    replayed on new batch contents of the same shape with eager work in
    between, reproduces the eager results every time."""
    c = local_case
    wl.setup_context(gpu_ctx, "box64", n_particles=24, n_rot=64)   # same seeds as local_case
    rng = np.random.default_rng(31)
    first = _batch(c, 24, rng, 300, 49)
    nR0, nT0 = first[1], first[3]
    same = []
    for _ in range(3):
        q = rng.normal(size=(int(nR0.sum()), 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        t = rng.uniform(-3, 3, size=(int(nT0.sum()), 2))
        same.append((q, nR0, t, nT0, rng.uniform(0.5, 1.5, int(nR0.sum())),
                     rng.uniform(0.5, 1.5, int(nT0.sum()))))
    batches = [first]
    outs = {}
    for mode in (1, 2, 0):
        gpu_ctx.set_local_options(mode, False)
        gpu_ctx.score_local(*batches[0], precision=thx.FP32)
        outs[mode] = gpu_ctx.get_local()
        assert gpu_ctx.local_mode() in (1, 2)
    for a, b in zip(outs[1], outs[2]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(outs[1], outs[0]):
        np.testing.assert_array_equal(a, b)
    # the 16-bit forms at both gather depths and launch forms: bit-identical
    for prec in (thx.BF16, thx.FP16):
        res = {}
        for ns in (4, 8):
            for mode in (1, 2):
                gpu_ctx.set_local_gather(ns)
                gpu_ctx.set_local_options(mode, False)
                gpu_ctx.score_local(*batches[0], precision=prec)
                res[(ns, mode)] = gpu_ctx.get_local()
                assert gpu_ctx.local_shape() == (mode, ns)
        for k in res:
            for a, b in zip(res[k], res[(4, 1)]):
                np.testing.assert_array_equal(a, b, err_msg=str(k))
    gpu_ctx.set_local_gather(0)
    for prec in (thx.BF16, thx.FP16):
        eager = []
        gpu_ctx.set_local_options(0, False)
        for b in same:
            gpu_ctx.score_local(*b, precision=prec)
            eager.append(gpu_ctx.get_local())
        gpu_ctx.set_local_options(0, True)
        for rep in range(2):
            for i, b in enumerate(same):
                gpu_ctx.score_global(0, thx.FP32)      # eager work between replays
                gpu_ctx.sync()
                gpu_ctx.score_local(*b, precision=prec)
                got = gpu_ctx.get_local()
                for a, e in zip(got, eager[i]):
                    np.testing.assert_array_equal(a, e)
    gpu_ctx.set_local_options(0, False)
