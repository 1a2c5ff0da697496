"""Particle filter of the local search on the GPU (needs gfx950).

Stage kernels against the oracle on the same samples:
  * inferACG matrix, mean, calVari concentrations and balance weights: the
    same fixed-point iteration count and agreement to 1e-9 (FP64 on both
    sides; only the summation order differs);
  * shift means, deviations and balance weights to 1e-12 relative;
  * peak factor, half-height peak and top index exactly.
The whole local search (perturb, balance, score, peak, calVari, resample,
convergence) is checked by what it must achieve on synthetic particles: from
a global scan on a 10-degree grid it converges and refines the poses below
the grid's spacing.
"""
import numpy as np
import pytest

import oracle
import thx
import workloads as wl

pytestmark = pytest.mark.gpu


def _qmat(m):
    w, x, y, z = m
    return np.array([[w, -x, -y, -z], [x, w, -z, y], [y, z, w, -x], [z, -y, x, w]])


def _acg_sets(rng, P, n):
    out = np.zeros((P, n, 4))
    for p in range(P):
        lam = np.r_[1.0, rng.uniform(0.005, 0.5, 3)]
        m = rng.normal(size=4)
        m /= np.linalg.norm(m)
        v = rng.normal(size=(n, 4)) * np.sqrt(lam)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        out[p] = v @ _qmat(m).T
    return out


@pytest.mark.parametrize("n", [40, 300, 2000])
def test_rot_stats_match_oracle(gpu_ctx, n):
    rng = np.random.default_rng(n)
    q = _acg_sets(rng, 6, n)
    A, mean, k, w, it = gpu_ctx.pf_rot_stats(q)
    for p in range(6):
        A0, it0 = oracle.infer_acg(q[p])
        assert it[p] == it0, (p, it[p], it0)
        np.testing.assert_allclose(A[p], A0, atol=1e-9)
        m0, k0 = oracle.cal_vari_rot(q[p])
        assert abs(abs(mean[p] @ m0) - 1) < 1e-9
        np.testing.assert_allclose(k[p], k0, rtol=1e-8)
        np.testing.assert_allclose(w[p], oracle.balance_rot(q[p]), rtol=1e-9)


def test_trans_stats_match_oracle(gpu_ctx):
    rng = np.random.default_rng(5)
    t = rng.normal(0, 1.5, size=(7, 33, 2)) + rng.normal(0, 2, size=(7, 1, 2))
    m, s, w = gpu_ctx.pf_trans_stats(t)
    for p in range(7):
        m0, s0, w0 = oracle.trans_stats(t[p])
        np.testing.assert_allclose(m[p], m0, rtol=1e-12)
        np.testing.assert_allclose(s[p], s0, rtol=1e-12)
        np.testing.assert_allclose(w[p], w0, rtol=1e-10)


@pytest.mark.parametrize("set_factor", [False, True])
def test_keep_peak_matches_oracle(gpu_ctx, set_factor):
    rng = np.random.default_rng(9)
    u = np.exp(rng.normal(0, 3, size=(5, 777)))
    u[1, 10] = u[1, 20] = u[1].max() * 2          # tied maxima: the first one wins
    f_in = np.array([1e-3, 0.2, 0.5, 0.01, 0.3])
    out, f, top = gpu_ctx.pf_keep_peak(u, f_in, set_factor)
    for p in range(5):
        ref, f0, top0 = oracle.keep_peak(u[p], f_in[p], set_factor)
        assert f[p] == f0 and top[p] == top0
        np.testing.assert_array_equal(out[p], ref)


@pytest.mark.parametrize("prec", [thx.FP32, thx.FP16])
def test_local_search_refines_global_assignments(gpu_ctx, prec):
    """Global scan on a coarse grid, particle states from it, local phases to
    convergence (FP32 or FP16 cross term): the final poses must beat the
    grid's own resolution."""
    print("\n[setup] box64 P=64 R=1000 T=49 local search", flush=True)
    info = wl.setup_context(gpu_ctx, "box64", n_particles=64, n_rot=1000)
    quat, trans = info["quat"], info["trans"]
    step = wl.sampling_step(quat)
    gpu_ctx.score_global(0, thx.FP32)
    gR, gT, _ = gpu_ctx.get_best(0)
    grid_err = wl.geodesic_angle(quat[gR], info["true_quat"])
    ms = wl.WORKLOADS["box64"]["n_rot"]
    gpu_ctx.pf_from_global(0, 300, 25, ms ** (-1.0 / 3), 0.3, 0.5, 1.0, seed=3)
    st0 = gpu_ctx.pf_get_state()
    assert np.all(st0["k"] > 0) and np.all(st0["s01"] > 0)
    assert np.all((st0["peakR"] >= 1e-3) & (st0["peakR"] <= 0.5))
    ran = gpu_ctx.pf_refine(None, prec, after_global=True, max_phase=40, seed=4)
    st = gpu_ctx.pf_get_state()
    err = wl.geodesic_angle(st["topR"], info["true_quat"])
    conv = st["nPhase"] >= 0
    terr = np.linalg.norm(st["topT"] - info["true_shift"], axis=1)
    print(f"phases {ran}, converged {conv.mean():.2f}, grid error median "
          f"{np.degrees(np.median(grid_err)):.2f} deg, refined median "
          f"{np.degrees(np.median(err)):.2f} deg (step {np.degrees(step):.1f}), shift error "
          f"median {np.median(terr):.2f} px")
    assert conv.mean() >= 0.9
    assert np.all(st["nPhase"][conv] >= 10)
    assert np.median(err) < 0.5 * np.median(grid_err)
    assert np.median(terr) < 0.5
    # weights are normalised per particle, samples stay unit quaternions
    np.testing.assert_allclose(st["wR"].sum(1), 1.0, rtol=1e-9)
    np.testing.assert_allclose(np.linalg.norm(st["quat"], axis=2), 1.0, rtol=1e-9)


def test_class_draw_and_per_class_states(gpu_ctx):
    """Two classes (independently seeded box-64 maps, 24 particles of each):
    the drawn class of every particle is among the classes that survive the
    class half-height peak (marginal >= 0.99 of the particle's largest, the
    merged marginals of both scans on one baseline); >= 90% of particles get
    the class they were made from; and each particle's state equals, bit for
    bit, the single-class update from its drawn class with the same seed."""
    w = wl.WORKLOADS["box64"]
    sd = wl.seeds("box64")
    iCol, iRow, iSig = wl.pixel_set(w["box"], w["cutoff"])
    gpu_ctx.set_pixels(iCol, iRow, iSig, w["box"], w["pf"])
    per, parts = 24, []
    for c in range(2):
        gpu_ctx.synth_volume(c, w["box"], w["pf"], w["n_blob"], w["blob_sigma"],
                             w["blob_radius"], sd["volume"] + 17 * c)
        gpu_ctx.synth_particles(c, per, w["max_shift"], w["df_min"], w["df_max"], w["snr"],
                                sd["particles"] + c, **wl.CTF_DEFAULTS)
        parts.append(gpu_ctx.get_particles())
    gpu_ctx.set_particles(*[np.concatenate([x[i] for x in parts]) for i in range(3)])
    truth = np.repeat([0, 1], per)
    quat, trans = wl.grid("box64", 1000)
    gpu_ctx.set_grid(quat, trans)
    for c in range(2):
        gpu_ctx.score_global(c, thx.FP32)
    wC, _, _, _ = gpu_ctx.get_global(2)
    args = (300, 25, 1000 ** (-1.0 / 3), 0.3, 0.5, 1.0)
    cls = gpu_ctx.pf_from_global_classes(2, *args, seed=3)
    surv = wC >= 0.99 * wC.max(1, keepdims=True)
    assert np.all(surv[np.arange(2 * per), cls])
    print(f"drawn class = true class for {np.mean(cls == truth):.3f}")
    assert np.mean(cls == truth) >= 0.9
    st = gpu_ctx.pf_get_state()
    for c in range(2):
        gpu_ctx.pf_from_global(c, *args, seed=3)
        ref = gpu_ctx.pf_get_state()
        m = cls == c
        for key in ("quat", "trans", "wR", "wT", "k", "s01", "peakR"):
            np.testing.assert_array_equal(st[key][m], ref[key][m])
    # the local search takes the drawn classes
    gpu_ctx.pf_from_global_classes(2, *args, seed=3)
    ran = gpu_ctx.pf_refine(cls, thx.BF16, after_global=True, max_phase=12, seed=4)
    assert ran >= 1


def _two_class_context(ctx, per=24, snr=None):
    """Two independently seeded box-64 maps, `per` particles made from each
    (at the workload's SNR unless given), the 1,000-rotation grid, and an FP32
    scan of both classes."""
    w = wl.WORKLOADS["box64"]
    sd = wl.seeds("box64")
    iCol, iRow, iSig = wl.pixel_set(w["box"], w["cutoff"])
    ctx.set_pixels(iCol, iRow, iSig, w["box"], w["pf"])
    parts = []
    for c in range(2):
        ctx.synth_volume(c, w["box"], w["pf"], w["n_blob"], w["blob_sigma"], w["blob_radius"],
                         sd["volume"] + 17 * c)
        ctx.synth_particles(c, per, w["max_shift"], w["df_min"], w["df_max"],
                            w["snr"] if snr is None else snr, sd["particles"] + c,
                            **wl.CTF_DEFAULTS)
        parts.append(ctx.get_particles())
    ctx.set_particles(*[np.concatenate([x[i] for x in parts]) for i in range(3)])
    quat, trans = wl.grid("box64", 1000)
    ctx.set_grid(quat, trans)
    for c in range(2):
        ctx.score_global(c, thx.FP32)
    return quat, trans


@pytest.mark.parametrize("n_cls", [1, 2])
def test_post_scan_matches_oracle(gpu_ctx, n_cls):
    """The whole post-scan update (src/Optimiser.cpp:926-1084: class draw,
    rotation peak factor and half-height peak, resampling of 300 rotations and
    25 shifts from the grid, calVari, scanning minima) against the oracle's
    restatement with the same random draws: drawn classes and resampled grid
    samples exactly, weights to 1e-12, peak factors to 1e-14, shift deviations
    to 1e-12 and concentrations to 1e-8 (FP64 inferACG, summation order
    differs) wherever the drawn rotations make calVari well posed."""
    print(f"\n[post-scan] box64 two classes, {n_cls} used", flush=True)
    # at SNR 0.003 the marginals are broad enough that most particles draw at
    # least four independent rotations
    quat, trans = _two_class_context(gpu_ctx, snr=0.003)
    P, R, T = gpu_ctx.P, len(quat), len(trans)
    # one class: its own marginals exactly; two: both on one baseline (the
    # class draw's u_k; each class's rows then carry a common factor, which the
    # peak keeping and resampling do not see beyond rounding)
    wC, wR, wT, base = gpu_ctx.get_global(n_cls)
    rng = np.random.default_rng(40 + n_cls)
    mLR, mLT = 300, 25
    u0C, pickC = rng.uniform(0, 1, P) / n_cls, rng.uniform(0, 1, P)
    u0R, u0T = rng.uniform(0, 1, P) / mLR, rng.uniform(0, 1, P) / mLT
    min_r, min_t, pfS = 0.1, 0.3, 0.5
    cls = gpu_ctx.pf_from_global_given(n_cls, mLR, mLT, min_r, min_t, pfS, 1.0, u0R, u0T,
                                       u0C if n_cls > 1 else None,
                                       pickC if n_cls > 1 else None)
    st = gpu_ctx.pf_get_state()
    pR, pT = np.full(R, 1.0 / R), np.full(T, 1.0 / T)
    kmin, compared = (min_r / pfS) ** 2, 0
    for p in range(P):
        ref = oracle.post_scan(wC[p, :n_cls], np.zeros(n_cls), wR[p, :n_cls], wT[p, :n_cls],
                               pR, pT, quat, trans, mLR, mLT, u0C[p], pickC[p], u0R[p], u0T[p],
                               min_r, min_t, pfS)
        assert cls[p] == ref["cls"], p
        np.testing.assert_array_equal(st["quat"][p], ref["q"])
        np.testing.assert_allclose(st["wR"][p], ref["wR"], rtol=1e-12)
        np.testing.assert_array_equal(st["trans"][p], ref["t"])
        np.testing.assert_allclose(st["wT"][p], ref["wT"], rtol=1e-12)
        assert st["peakR"][p] == pytest.approx(ref["peakR"], rel=1e-14)
        np.testing.assert_allclose(st["s01"][p], ref["s"], rtol=1e-12)
        # calVari's inferACG is defined only when the drawn rotations span the
        # four dimensions of their scatter matrix; with fewer than four
        # independent samples (a sharp peak leaves 1-3 grid rotations) its
        # fixed point inverts a singular matrix on both sides and only the
        # scanning minimum is certain
        sv = np.linalg.svd(ref["q"], compute_uv=False)
        if sv[-1] > 1e-6 * sv[0]:
            np.testing.assert_allclose(st["k"][p], ref["k"], rtol=1e-8)
            compared += 1
        else:
            assert np.all(st["k"][p] >= kmin) and np.all(ref["k"] >= kmin)
    print(f"concentrations compared for {compared} of {P} particles (the others drew fewer "
          f"than four independent rotations)")
    assert compared >= P // 4
    if n_cls > 1:
        print(f"classes drawn: {np.bincount(cls, minlength=2)}")
        assert len(set(cls.tolist())) == 2


def test_converge_matches_oracle(gpu_ctx):
    """k_converge over scripted per-phase concentrations and deviations against
    the oracle's convergence rule: the same stopping phase for every particle,
    after a global scan (checks from phase 10) and in a local search (from 3)."""
    rng = np.random.default_rng(77)
    P, nPh = 256, 40
    # variances that shrink at a random rate, then level off at a random phase,
    # with noise: every particle stops somewhere, many at distinct phases
    rate = rng.uniform(0.85, 1.0, size=(P, 1))
    stop = rng.integers(5, 35, size=(P, 1))
    ph = np.arange(1, nPh + 1)[None, :]
    lev = rate ** np.minimum(ph, stop) * rng.uniform(0.97, 1.03, size=(P, nPh))
    k = np.stack([lev ** 2 * rng.uniform(0.5, 2, size=(P, 1)) for _ in range(3)], axis=2)
    s01 = np.stack([np.sqrt(lev) * rng.uniform(0.5, 2, size=(P, 1)) for _ in range(2)], axis=2)
    k, s01 = k.transpose(1, 0, 2), s01.transpose(1, 0, 2)      # [nPh][P][.]
    for after_global, phase0, min_phase in ((True, 1, 10), (False, 0, 3)):
        got = gpu_ctx.pf_converge(k, s01, after_global)
        ref = np.array([oracle.converge(k[:, p], s01[:, p], phase0=phase0, min_phase=min_phase)
                        for p in range(P)])
        print(f"after_global={after_global}: stopping phases {np.bincount(ref + 1)[:15]}...")
        np.testing.assert_array_equal(got, ref)
        assert len(set(ref.tolist())) >= 10


def test_mat4_helpers_against_lapack(gpu_ctx):
    """The 4x4 inverse / determinant and the top eigenvector used by the
    inferACG stages, against numpy's LAPACK routines (independent of the
    oracle's own 4x4 code): general and ill-conditioned symmetric positive
    definite matrices, and a singular one."""
    rng = np.random.default_rng(8)
    n = 200
    A = rng.normal(size=(n, 4, 4))
    spd = []
    for i in range(n):
        Q, _ = np.linalg.qr(rng.normal(size=(4, 4)))
        ev = np.sort(10.0 ** rng.uniform(-4, 0, 4))[::-1]
        ev[0] = 1.0 + rng.uniform(0.05, 1.0)          # a separated top eigenvalue
        spd.append(Q @ np.diag(ev) @ Q.T)
    M = np.concatenate([A, np.array(spd), np.ones((1, 4, 4))])
    inv, det, top = gpu_ctx.pf_mat4(M)
    for i in range(len(M) - 1):
        ref = np.linalg.inv(M[i])
        np.testing.assert_allclose(inv[i], ref, rtol=0, atol=1e-9 * np.abs(ref).max())
        assert det[i] == pytest.approx(np.linalg.det(M[i]), rel=1e-9)
    assert det[-1] == 0.0 and np.all(inv[-1] == 0.0)
    for i in range(n, 2 * n):
        w, V = np.linalg.eigh(M[i])
        v = V[:, -1] * np.sign(V[np.argmax(np.abs(V[:, -1])), -1])
        np.testing.assert_allclose(top[i], v, atol=1e-10)


def test_refine_phase_graph_replays_equal_eager(gpu_ctx):
    """The local search's phases as one captured graph per sample-buffer
    assignment, replayed for every phase (the work list, perturbation scale
    and random-stream phases come from device memory): the final states equal
    the eager run's bit for bit, for both launch shapes of the scorer, and on
    a second search that reuses the cached graphs after eager work on the
    same stream."""
    print("\n[setup] box64 P=64 R=1000 T=49 refine: graph vs eager", flush=True)
    wl.setup_context(gpu_ctx, "box64", n_particles=64, n_rot=1000)
    gpu_ctx.score_global(0, thx.FP32)
    ms = wl.WORKLOADS["box64"]["n_rot"]
    args = (0, 300, 25, ms ** (-1.0 / 3), 0.3, 0.5, 1.0)
    keys = ("quat", "trans", "wR", "wT", "k", "s01", "topR", "topT", "peakR", "nPhase")

    def run(mode, graph):
        gpu_ctx.set_local_options(mode, graph)
        gpu_ctx.pf_from_global(*args, seed=3)
        ran = gpu_ctx.pf_refine(None, thx.BF16, after_global=True, max_phase=14, seed=4)
        return ran, gpu_ctx.pf_get_state()

    try:
        for mode in (1, 2):
            ran0, ref = run(mode, False)
            assert ran0 >= 10
            for rep in range(2):
                ran, st = run(mode, True)
                assert ran == ran0
                for k in keys:
                    np.testing.assert_array_equal(st[k], ref[k], err_msg=f"{k} mode {mode} rep {rep}")
                gpu_ctx.score_global(0, thx.FP32)   # eager work between the replays
                gpu_ctx.sync()
    finally:
        gpu_ctx.set_local_options(0, False)
