"""Particle images to packed particle data on the GPU (needs gfx950): the
image path of Optimiser::initImg (substractBgImg + fwImg,
src/Optimiser.cpp:4917-4951, 5003-5013) and allocPreCal's datP packing
(src/Optimiser.cpp:8005-8016), against the FP64 numpy oracle on the same
images.

Images are seeded: a Gaussian blob off centre on a background of known mean
and spread, so every image has its own background statistics.  The GPU
computes the statistics in FP64 and the FFT in FP32: every packed coefficient
must agree with the oracle to 1e-5 of the largest magnitude, at box 64 and
128 with their pixel sets, at a size with factors 3 and 5 (box 160), and for
a background radius that reaches past the edge midpoints (only the corners
are background); each case unmasked and with the soft mask of radius equal to
the background radius (6-pixel cosine edge) applied.  Boxes up to 256 with N a multiple of 4 hold each image in
registers (three register budgets: box 64 / 128, box 100 / 160 / 180, and box 192 / 256
with 512-thread blocks); box 66 (not a multiple of 4) takes the generic
kernel.  Boxes 100 and 180 (N / 2 not a multiple of 4) move one group of
four pixels per row across the row's wrap.  The CTF and sigma arrays pass through unchanged.  Data set
this way scores exactly like the same data uploaded with set_particles.
"""
import numpy as np
import pytest

import oracle
import thx
import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ictx(gpu_ctx):
    """A context of its own: pixel sets of several sizes."""
    ctx = thx.Context(0)
    yield ctx
    ctx.close()


def _images(P, N, seed):
    rng = np.random.default_rng(seed)
    h = N // 2
    yy, xx = np.mgrid[0:N, 0:N]
    out = np.empty((P, N, N), np.float32)
    for p in range(P):
        cx, cy = h + rng.uniform(-3, 3), h + rng.uniform(-3, 3)
        blob = 4.0 * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * (N / 10) ** 2))
        out[p] = blob + rng.normal(rng.uniform(-2, 2), rng.uniform(0.5, 3), size=(N, N))
    return out


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("N,cutoff,rbg", [(64, 24, 30.0), (128, 50, 60.0), (160, 64, 76.0),
                                          (64, 24, 40.0), (66, 24, 31.0), (192, 75, 90.0),
                                          (256, 100, 120.0), (100, 40, 48.0),
                                          (180, 70, 85.0)])
def test_particle_images_match_oracle(ictx, N, cutoff, rbg, masked):
    P = 48
    print(f"\n[images] N={N} cutoff={cutoff} bg radius={rbg} P={P} masked={masked}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, cutoff)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    imgs = _images(P, N, 7 + N)
    rng = np.random.default_rng(N)
    ctf = rng.uniform(-1, 1, size=(P, K)).astype(np.float32)
    sig = rng.uniform(0.5, 2, size=(P, K)).astype(np.float32)
    ms = ictx.set_particle_images(imgs, rbg, ctf, sig, soft_mask=masked)
    dat, ctf2, sig2 = ictx.get_particles()
    got = dat[..., 0] + 1j * dat[..., 1]
    ref = oracle.particle_images(imgs, rbg, iCol, iRow, masked=masked)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"max error {err:.2e} of the largest coefficient; times (ms) {ms}")
    assert err <= 1e-5
    np.testing.assert_array_equal(ctf2, ctf)
    np.testing.assert_array_equal(sig2, sig)


def test_images_score_as_uploaded_data(ictx):
    """The packed data scores exactly as the same arrays given to
    set_particles (both paths leave identical operands)."""
    N, P = 64, 16
    print(f"\n[images] scoring check N={N} P={P}", flush=True)
    info_ctx = thx.Context(0)
    try:
        info = wl.setup_context(info_ctx, "box64", n_particles=4, n_rot=64)
        vol = info_ctx.get_volume(0, 128)
    finally:
        info_ctx.close()
    iCol, iRow, iSig = info["iCol"], info["iRow"], info["iSig"]
    K = len(iCol)
    imgs = _images(P, N, 99)
    rng = np.random.default_rng(5)
    ctf = rng.uniform(-1, 1, size=(P, K)).astype(np.float32)
    sig = rng.uniform(0.5, 2, size=(P, K)).astype(np.float32)
    results = []
    for mode in ("images", "arrays"):
        ictx.set_pixels(iCol, iRow, iSig, N, 2)
        ictx.set_volume(0, vol, 128)
        if mode == "images":
            ictx.set_particle_images(imgs, 30.0, ctf, sig)
            dat = ictx.get_particles()[0]
        else:
            ictx.set_particles(dat, ctf, sig)
        ictx.set_grid(info["quat"], info["trans"])
        ictx.score_global(0, thx.FP32)
        results.append(ictx.get_best(0))
    for a, b in zip(*results):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("N,cutoff,rbg,P", [(64, 24, 30.0, 48), (128, 50, 60.0, 300)])
def test_init_sigma_matches_oracle(ictx, N, cutoff, rbg, P):
    """Initial sigma^2 and sigRcp per shell from the last images
    (Optimiser::initSigma) against the FP64 oracle on the same images: the
    GPU sums FP32 spectra in FP64, so every shell agrees to 1e-4 relative;
    with apply the particles' sigRcp becomes the shell value at iSig."""
    print(f"\n[sigma] N={N} P={P}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, cutoff)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    imgs = _images(P, N, 3 + N)
    ones = np.ones((P, K), np.float32)
    # masked particle data; the estimate still reads the unmasked spectra
    ictx.set_particle_images(imgs, rbg, ones, ones, soft_mask=True)
    nS = N // 2 - 1
    with pytest.raises(thx.THXError):
        ictx.init_sigma(nS + 1)
    sg, rcp, ms = ictx.init_sigma(nS, apply=True)
    ref, ref_rcp = oracle.init_sigma(imgs, rbg, nS)
    err = np.max(np.abs(sg - ref) / np.abs(ref))
    print(f"max relative error {err:.2e}; kernels {ms:.3f} ms")
    assert err <= 1e-4
    np.testing.assert_allclose(rcp, -0.5 / sg, rtol=1e-12)
    got = ictx.get_particles()[2]
    want = np.broadcast_to(ref_rcp[np.asarray(iSig)].astype(np.float32), (P, K))
    np.testing.assert_allclose(got, want, rtol=1e-4)


def test_init_sigma_apply_needs_image_particles(ictx):
    """After set_particles replaces the particle arrays (same count), the
    spectra of the last images no longer describe them: apply is refused,
    the estimate alone is still available."""
    N, P = 64, 16
    print(f"\n[sigma] apply after set_particles N={N} P={P}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    ones = np.ones((P, K), np.float32)
    ictx.set_particle_images(_images(P, N, 11), 30.0, ones, ones)
    ictx.set_particles(np.zeros((P, K, 2), np.float32), ones, ones)
    with pytest.raises(thx.THXError, match="not the last images"):
        ictx.init_sigma(N // 2 - 1, apply=True)
    ictx.init_sigma(N // 2 - 1, apply=False)


def test_set_particles_rejects_wrong_shapes(ictx):
    """ctf / sig of another shape than (P, K) raise before the library reads
    the host arrays."""
    N, P = 64, 4
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    with pytest.raises(ValueError):
        ictx.set_particles(np.zeros((P, K, 2), np.float32), np.ones((P, K)), np.ones((P, K - 1)))
    with pytest.raises(ValueError):
        ictx.set_particle_images(_images(P, N, 1), 30.0, np.ones((P, K - 3)), np.ones((P, K)))


def test_init_sigma_needs_images(gpu_ctx):
    """Without particle images there is nothing to estimate from: a clear
    error, not a result."""
    ctx = thx.Context(0)
    try:
        with pytest.raises(thx.THXError, match="no particle images"):
            ctx.init_sigma(10)
    finally:
        ctx.close()


def test_more_images_than_a_y_grid(ictx):
    """70,000 images of 16^2 (more than 65,535; the kernels put particles on
    the grid's x dimension): every particle packed, soft-masked, and given its
    sigRcp."""
    N, P = 16, 70_000
    print(f"\n[images] N={N} P={P} masked, with sigma", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, 6)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    rng = np.random.default_rng(70)
    imgs = rng.normal(0.0, 1.0, size=(P, N, N)).astype(np.float32)
    ones = np.ones((P, K), np.float32)
    ictx.set_particle_images(imgs, 5.0, ones, ones, soft_mask=True)
    sg, rcp, _ = ictx.init_sigma(N // 2 - 1, apply=True)
    dat, _, sig = ictx.get_particles()
    got = dat[..., 0] + 1j * dat[..., 1]
    ref = oracle.particle_images(imgs, 5.0, iCol, iRow, masked=True)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"max error {err:.2e} of the largest coefficient", flush=True)
    assert err <= 1e-5
    np.testing.assert_array_equal(sig, np.broadcast_to(rcp[np.asarray(iSig)].astype(np.float32),
                                                       (P, K)))


@pytest.mark.parametrize("name", ["box64", "box128", "box160", "box192", "box256"])
def test_ctf_from_attributes_matches_oracle(ictx, name):
    """The particles' CTF evaluated on the device from their CTF attributes
    (thx_ctx_set_ctf: CTF(RFLOAT*), src/CTF.cpp:109-145, per particle as
    allocPreCal packs ctfP on the fly) at every workload's pixel set, against
    the oracle's FP32 restatement: per-particle voltage, defocus, astigmatism
    angle, Cs, amplitude contrast and phase shift.  Both sides round the
    phase ki to float (up to ~2,000 rad at box 256), so the bar is
    4 ulp of the largest |ki| plus 2e-5."""
    w = wl.WORKLOADS[name]
    N, P = w["box"], 12
    print(f"\n[ctf] {name}: N={N} cutoff={w['cutoff']} P={P}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, w["cutoff"])
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    K = len(iCol)
    rng = np.random.default_rng(N)
    attr = np.stack([rng.uniform(200e3, 300e3, P), rng.uniform(5e3, 40e3, P),
                     rng.uniform(5e3, 40e3, P), rng.uniform(0, np.pi, P),
                     rng.uniform(2.0e7, 2.7e7, P), rng.uniform(0.05, 0.15, P),
                     np.where(np.arange(P) % 3 == 0, rng.uniform(0, 1.5, P), 0.0)], 1)
    ictx.set_particles(np.zeros((P, K, 2), np.float32), None, np.ones((P, K), np.float32))
    check_score = name == "box64"
    if check_score:
        # no CTF yet: scoring is refused, not run on undefined values
        ictx.set_volume(0, np.zeros((2 * N, 2 * N, N + 1, 2), np.float32), 2 * N)
        ictx.set_grid(wl.super_fibonacci(8), [[0.0, 0.0]])
        with pytest.raises(thx.THXError, match="no CTF"):
            ictx.score_global(0, thx.FP32)
    px = 1.5
    ictx.set_ctf(attr.astype(np.float32), px)
    got = ictx.get_particles()[1]
    worst = 0.0
    for p in range(P):
        a = attr[p].astype(np.float32)
        ref = oracle.ctf(iCol, iRow, N, px, a[0], a[1], a[2], a[3], a[4], a[5], a[6])
        lam = 12.2643247 / np.sqrt(a[0] * (1 + a[0] * 0.978466e-6))
        u2 = (np.asarray(iCol) ** 2 + np.asarray(iRow) ** 2) / (px * N) ** 2
        kmax = np.max(np.pi * lam * max(a[1], a[2]) * u2 + np.pi / 2 * a[4] * lam ** 3 * u2 ** 2)
        tol = 2e-5 + 4 * kmax * 2.0 ** -23
        err = np.abs(got[p] - ref).max()
        worst = max(worst, err / tol)
        assert err <= tol, (p, err, tol)
    print(f"worst error / bar {worst:.2f}")
    if check_score:
        ictx.score_global(0, thx.FP32)   # now set: scoring runs


@pytest.mark.parametrize("N,rbg", [(64, 30.0), (128, 60.0), (66, 31.0)])
def test_image_stats_match_oracle(ictx, N, rbg):
    """statImg's mean background deviation of the background-subtracted
    images and normaliseImg's scale (src/Optimiser.cpp:4796-4895,
    4987-5001) against the oracle: stdN to 1e-6 (the GPU sums the squares of
    the FP32 images it writes, the oracle works in FP64), scale = 1 / stdN
    rounded to float, and the packed data carry it (the image comparisons
    above include it)."""
    P = 40
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    imgs = _images(P, N, 5 + N)
    ictx.set_particle_images(imgs, rbg, None, None)
    stdN, scale = ictx.image_stats()
    y, bg = oracle.subtract_bg(imgs, rbg)
    ref = oracle.stat_img(y, bg)
    print(f"\n[stats] N={N}: stdN {stdN:.9f} (oracle {ref:.9f}), scale {scale:.9f}")
    assert abs(stdN - ref) <= 1e-6
    assert scale == np.float32(1.0 / stdN)


def test_load_thu_equals_direct_upload(ictx, tmp_path):
    """Particles read from a .thu file and its MRC stacks (thx_ctx_load_thu:
    "index@stack" paths under a prefix, as Optimiser::initImg reads them,
    CTF from the file's attributes) equal, bit for bit, the same images and
    attributes given to set_particle_images + set_ctf; a sub-range of the file
    likewise."""
    import formats
    N, P, rbg, px = 64, 24, 30.0, 1.5
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    imgs = _images(P, N, 123)
    rng = np.random.default_rng(8)
    attr = np.stack([np.full(P, 300e3), rng.uniform(1e4, 3e4, P), rng.uniform(1e4, 3e4, P),
                     rng.uniform(0, 3, P), np.full(P, 2.7e7), np.full(P, 0.07),
                     np.zeros(P)], 1)
    attr = np.array([[float(f"{x:18.6f}") for x in a] for a in attr], np.float32)
    d = tmp_path / "particles"
    d.mkdir()
    thx.mrc_write(str(d / "a.mrcs"), imgs[:10], px)
    thx.mrc_write(str(d / "b.mrcs"), imgs[10:], px)
    thu = str(tmp_path / "meta.thu")
    with open(thu, "w") as fh:
        for p in range(P):
            name = f"{p + 1}@a.mrcs" if p < 10 else f"{p - 9}@b.mrcs"
            fh.write(formats.thu_line_input(attr[p], name, "mic.mrc", (0.0, 0.0), 1))
    for first, count in ((0, P), (5, 12)):
        print(f"\n[thu] particles {first}..{first + count - 1}", flush=True)
        ictx.load_thu(thu, str(d) + "/", first, count, px, rbg, soft_mask=True)
        dat1, ctf1, _ = ictx.get_particles()
        sl = slice(first, first + count)
        ictx.set_particle_images(imgs[sl], rbg, None, None, soft_mask=True)
        ictx.set_ctf(attr[sl], px)
        dat2, ctf2, _ = ictx.get_particles()
        np.testing.assert_array_equal(dat1, dat2)
        np.testing.assert_array_equal(ctf1, ctf2)
    with pytest.raises(thx.THXError, match="has 24 particles"):
        ictx.load_thu(thu, str(d) + "/", 20, 10, px, rbg)


def test_refresh_sigma_matches_oracle(ictx):
    """The sigma refresh from best poses (Optimiser::allReduceSigma) against
    the FP64 oracle: two classes, two groups, masked and unmasked spectra of
    real-space images, per-particle CTF attributes, poses, shifts and
    recentring offsets.  The GPU forms the model in FP32 (trilinear slice, CTF,
    phase) from FP32 spectra: every shell's sigma^2 within 2e-4 relative; with
    apply, every particle's sigRcp is its group's value at its pixel's shell."""
    N, P, rbg, px, r = 64, 40, 30.0, 1.5, 20
    nS = N // 2 - 1
    w = wl.WORKLOADS["box64"]
    print(f"\n[sigma refresh] N={N} P={P} r={r} shells={nS}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    vols = []
    for c in range(2):
        ictx.synth_volume(c, N, 2, w["n_blob"], w["blob_sigma"], w["blob_radius"], 40 + c)
        vols.append(ictx.get_volume(c, 2 * N))
    imgs = _images(P, N, 77)
    ictx.set_particle_images(imgs, rbg, None, None, soft_mask=True)
    rng = np.random.default_rng(12)
    attr = np.stack([np.full(P, 300e3), rng.uniform(1e4, 3e4, P), rng.uniform(1e4, 3e4, P),
                     rng.uniform(0, 3, P), np.full(P, 2.7e7), np.full(P, 0.07),
                     np.where(np.arange(P) % 4 == 0, 0.3, 0.0)], 1).astype(np.float32)
    ictx.set_ctf(attr, px)
    quat = rng.normal(size=(P, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    trans = rng.uniform(-3, 3, size=(P, 2))
    offset = rng.uniform(-1, 1, size=(P, 2))
    cls = rng.integers(0, 2, P).astype(np.int32)
    grp = (np.arange(P) % 2).astype(np.int32)
    mask_r = 40.0
    sig, rcp, ms = ictx.refresh_sigma(quat, trans, r, nS, mask_r, px, cls=cls, offset=offset,
                                      group=grp, n_group=2, apply=True)
    specM = oracle.image_spectra(imgs, rbg, masked=True)
    specN = oracle.image_spectra(imgs, rbg, masked=False)
    ref, ref_rcp = oracle.refresh_sigma(vols, 2 * N, 2, specM, specN, attr, px, cls, quat,
                                        trans, offset, grp, 2, r, nS, mask_r)
    err = np.abs(sig - ref) / np.abs(ref)
    print(f"max relative error {err.max():.2e}; kernels {ms:.3f} ms")
    assert err.max() <= 2e-4
    np.testing.assert_allclose(rcp, -0.5 / sig, rtol=1e-12)
    got = ictx.get_particles()[2]
    want = rcp[grp][:, np.asarray(iSig)].astype(np.float32)
    np.testing.assert_array_equal(got, want)


def test_insert_matches_oracle(ictx):
    """The reconstruction insert (Optimiser::reconstructRef's loop with
    Reconstructor::insertP, trilinear) against the FP64 oracle: two classes,
    two samples per particle, per-particle weights, shifts, recentring offsets
    and CTF attributes, the unmasked spectra of real-space images.  F and T
    within 1e-5 of their largest magnitude (the GPU forms each value in FP32
    as the reference's RFLOAT and accumulates in FP64), the insertDir sums to
    1e-12, and a second call without clearing doubles them."""
    N, P, m, rbg, px, rU = 64, 12, 2, 30.0, 1.5, 20
    V = 2 * N
    print(f"\n[insert] N={N} P={P} samples={m} rU={rU}", flush=True)
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    imgs = _images(P, N, 41)
    ictx.set_particle_images(imgs, rbg, None, None, soft_mask=True)
    rng = np.random.default_rng(6)
    attr = np.stack([np.full(P, 300e3), rng.uniform(1e4, 3e4, P), rng.uniform(1e4, 3e4, P),
                     rng.uniform(0, 3, P), np.full(P, 2.7e7), np.full(P, 0.07),
                     np.zeros(P)], 1).astype(np.float32)
    ictx.set_ctf(attr, px)
    quat = rng.normal(size=(P, m, 4))
    quat /= np.linalg.norm(quat, axis=2, keepdims=True)
    trans = rng.uniform(-3, 3, size=(P, m, 2))
    offset = rng.uniform(-1, 1, size=(P, 2))
    cls = rng.integers(0, 2, size=(P, m)).astype(np.int32)
    w = rng.uniform(0.5, 1.0, P)
    ms = ictx.insert(quat, trans, rU, px, cls=cls, offset=offset, w=w, n_cls=2)
    specN = oracle.image_spectra(imgs, rbg, masked=False)
    F0, T0, d0, n0 = oracle.reconstruct_insert(specN, attr, px, cls, quat, trans, offset, w / m,
                                               2, V, 2, rU)
    for k in range(2):
        F, T, d, n = ictx.get_reconstruction(k)
        eF = np.abs(F - F0[k]).max() / np.abs(F0[k]).max()
        eT = np.abs(T - T0[k]).max() / np.abs(T0[k]).max()
        print(f"class {k}: F error {eF:.2e}, T error {eT:.2e}, samples {n:.0f}; {ms:.3f} ms")
        assert eF <= 1e-5 and eT <= 1e-5
        np.testing.assert_allclose(d, d0[k], rtol=1e-12, atol=1e-12)
        assert n == n0[k]
    ictx.insert(quat, trans, rU, px, cls=cls, offset=offset, w=w, n_cls=2, clear=False)
    F, T, d, n = ictx.get_reconstruction(1)
    assert np.abs(F - 2 * F0[1]).max() <= 2e-5 * np.abs(F0[1]).max()
    assert n == 2 * n0[1]


def _model_images(ictx, N, P, seed, noise=0.3, n_cls=2, with_offset=True):
    """Real-space images whose spectra are each particle's best-pose model
    (CTF . shift(t - offset) . slice of its class, from the oracle's pinned
    pieces) plus white noise, centred as a stack holds them; the classes'
    volumes synthesised in the context.  Returns what the tests need."""
    w = wl.WORKLOADS["box64"]
    iCol, iRow, iSig = wl.pixel_set(N, 24)
    ictx.set_pixels(iCol, iRow, iSig, N, 2)
    vols = []
    for c in range(n_cls):
        ictx.synth_volume(c, N, 2, w["n_blob"], w["blob_sigma"], w["blob_radius"], 60 + c)
        vols.append(ictx.get_volume(c, 2 * N))
    rng = np.random.default_rng(seed)
    px, r = 1.5, N // 2 - 4
    attr = np.stack([np.full(P, 300e3), rng.uniform(1e4, 3e4, P), rng.uniform(1e4, 3e4, P),
                     rng.uniform(0, 3, P), np.full(P, 2.7e7), np.full(P, 0.07),
                     np.where(np.arange(P) % 5 == 0, 0.2, 0.0)], 1).astype(np.float32)
    quat = rng.normal(size=(P, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    trans = rng.uniform(-3, 3, size=(P, 2))
    offset = rng.uniform(-1, 1, size=(P, 2)) if with_offset else np.zeros((P, 2))
    cls = rng.integers(0, n_cls, P).astype(np.int32)
    i, j, row = oracle.spectrum_pixels(N)
    inr = i * i + j * j < r * r
    imgs = np.empty((P, N, N), np.float32)
    for l in range(P):
        c = oracle.ctf(i[inr], j[inr], N, px, *attr[l])
        t = trans[l] - offset[l]
        s = np.exp(-2j * np.pi * (i[inr] * t[0] + j[inr] * t[1]) / N)
        p = oracle.project(vols[cls[l]], 2 * N, oracle.rotmat(quat[l]), i[inr], j[inr], 2, f64=True)
        spec = np.zeros((N, N // 2 + 1), np.complex128)
        spec[row[inr], i[inr]] = c * s * (p[:, 0] + 1j * p[:, 1])
        img = np.fft.irfft2(spec, s=(N, N))
        img = np.fft.fftshift(img)
        imgs[l] = img + noise * img.std() * rng.normal(size=(N, N))
    return dict(vols=vols, attr=attr, quat=quat, trans=trans, offset=offset, cls=cls, imgs=imgs,
                px=px, r=r)


def test_refresh_scale_matches_oracle(ictx):
    """The intensity scale from best poses (Optimiser::refreshScale) against
    the FP64 oracle: images made from each particle's model plus noise, two
    classes, three groups, recentring offsets; then correctScale's two
    applications: the images and sigRcp (data / s, sigRcp x s^2, after which
    the scale refreshed again is 1 in every group) and the class volumes
    (x the first particle's group scale)."""
    N, P, rbg = 64, 45, 30.0
    print(f"\n[refresh scale] N={N} P={P}", flush=True)
    d = _model_images(ictx, N, P, 31)
    ictx.set_particle_images(d["imgs"], rbg, None, None, soft_mask=True)
    ictx.set_ctf(d["attr"], d["px"])
    ictx.init_sigma(N // 2 - 1, apply=True)
    grp = (np.arange(P) % 3).astype(np.int32)
    r, rS, rL = d["r"], 20, 1.5
    sc, ms = ictx.refresh_scale(d["quat"], d["trans"], r, rS, rL, d["px"], cls=d["cls"],
                                offset=d["offset"], group=grp, n_group=3)
    specN = oracle.image_spectra(d["imgs"], rbg, masked=False)
    ref = oracle.refresh_scale(d["vols"], 2 * N, 2, specN, d["attr"], d["px"], d["cls"],
                               d["quat"], d["trans"], d["offset"], grp, 3, r, rS, rL)
    print(f"scales {sc}, oracle {ref}; {ms:.3f} ms")
    np.testing.assert_allclose(sc, ref, rtol=1e-5)
    # without groups, and without coordinates (zero frequency of class 0)
    sc1, _ = ictx.refresh_scale(d["quat"], d["trans"], r, rS, rL, d["px"], cls=d["cls"],
                                offset=d["offset"], n_group=3)
    ref1 = oracle.refresh_scale(d["vols"], 2 * N, 2, specN, d["attr"], d["px"], d["cls"],
                                d["quat"], d["trans"], d["offset"], None, 3, r, rS, rL)
    np.testing.assert_allclose(sc1, ref1, rtol=1e-5)
    sc0, _ = ictx.refresh_scale(d["quat"], None, r, rS, rL, d["px"], group=grp, n_group=3,
                                coord=False)
    ref0 = oracle.refresh_scale(d["vols"], 2 * N, 2, specN, d["attr"], d["px"], None, d["quat"],
                                None, None, grp, 3, r, rS, rL, coord=False)
    np.testing.assert_allclose(sc0, ref0, rtol=1e-5)
    # correctScale(init = false)
    dat0, _, sig0 = ictx.get_particles()
    sc, _ = ictx.refresh_scale(d["quat"], d["trans"], r, rS, rL, d["px"], cls=d["cls"],
                               offset=d["offset"], group=grp, n_group=3, apply=1)
    dat1, _, sig1 = ictx.get_particles()
    f = (1.0 / sc[grp]).astype(np.float32)[:, None, None]
    np.testing.assert_allclose(dat1, dat0 * f, rtol=1e-6, atol=1e-6 * np.abs(dat0).max())
    np.testing.assert_allclose(sig1, sig0 * (sc[grp] ** 2).astype(np.float32)[:, None], rtol=1e-6)
    again, _ = ictx.refresh_scale(d["quat"], d["trans"], r, rS, rL, d["px"], cls=d["cls"],
                                  offset=d["offset"], group=grp, n_group=3)
    np.testing.assert_allclose(again, 1.0, rtol=1e-5)
    # correctScale(init = true): the volumes
    v0 = [ictx.get_volume(k, 2 * N) for k in range(2)]
    sc, _ = ictx.refresh_scale(d["quat"], d["trans"], r, rS, rL, d["px"], cls=d["cls"],
                               offset=d["offset"], group=grp, n_group=3, apply=2)
    for k in range(2):
        np.testing.assert_array_equal(ictx.get_volume(k, 2 * N),
                                      v0[k] * np.float32(sc[grp[0]]))


def test_norm_correction_matches_oracle(ictx):
    """The noise normalisation from best poses (Optimiser::normCorrection)
    against the FP64 oracle on the masked spectra: per-particle residual
    powers within 1e-4, the median (an exact selection on the device) equal
    to the oracle's median of the GPU's own powers, odd and even particle
    counts; applied, the packed data are multiplied by sqrt(m / norm) and a
    second pass sees the scaled spectra."""
    N, rbg = 64, 30.0
    for P in (33, 40):
        print(f"\n[norm correction] N={N} P={P}", flush=True)
        d = _model_images(ictx, N, P, 50 + P, noise=0.5)
        ictx.set_particle_images(d["imgs"], rbg, None, None, soft_mask=True)
        ictx.set_ctf(d["attr"], d["px"])
        r, rL, rN = d["r"], 2.0, 22.5
        # normCorrection projects at t, not t - offset: the model images carry
        # t - offset, so pass that as the pose's translation here
        t = d["trans"] - d["offset"]
        norm, m, ms = ictx.norm_correction(d["quat"], t, r, rL, rN, d["px"], cls=d["cls"])
        specM = oracle.image_spectra(d["imgs"], rbg, masked=True)
        rn, rm, rf = oracle.norm_correction(d["vols"], 2 * N, 2, specM, d["attr"], d["px"],
                                            d["cls"], d["quat"], t, r, rL, rN)
        err = np.abs(norm - rn) / rn
        print(f"norm max relative error {err.max():.2e}, median {m:.6g} vs {rm:.6g}; "
              f"{ms:.3f} ms")
        assert err.max() <= 1e-4
        assert m == oracle.gsl_median(norm)
        dat0 = ictx.get_particles()[0]
        norm, m, _ = ictx.norm_correction(d["quat"], t, r, rL, rN, d["px"], cls=d["cls"],
                                          apply=True)
        f = np.sqrt(m / norm).astype(np.float32)
        dat1 = ictx.get_particles()[0]
        np.testing.assert_allclose(dat1, dat0 * f[:, None, None], rtol=1e-6,
                                   atol=1e-6 * np.abs(dat0).max())
        norm2, _, _ = ictx.norm_correction(d["quat"], t, r, rL, rN, d["px"], cls=d["cls"])
        rn2, _, _ = oracle.norm_correction(d["vols"], 2 * N, 2, specM * f[:, None, None],
                                           d["attr"], d["px"], d["cls"], d["quat"], t, r, rL, rN)
        assert (np.abs(norm2 - rn2) / rn2).max() <= 1e-4
