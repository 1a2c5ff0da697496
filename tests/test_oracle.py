"""Pinning of the CPU oracle (no GPU).

The reference ships no golden vectors, tests or fixtures for the expectation
path, so each oracle function is pinned by checks that do not restate its own
formula: closed forms, textbook constants, independent algorithms (numpy FFTs,
a Hermitian-extended full-volume trilinear interpolator), invariants and
hand-worked cases.  Each test names the plausible mistake it would catch.
"""
import math

import numpy as np
import pytest

import oracle
import workloads as wl


# ---------------------------------------------------------------------------
# a1 pixel set (Optimiser::allocPreCalIdx)
# ---------------------------------------------------------------------------

def _brute_pixels(N, rU, rL=0.0):
    i, j = np.meshgrid(np.arange(0, N // 2 + 1), np.arange(-N // 2, N // 2), indexing="xy")
    i, j = i.ravel(), j.ravel()
    r2 = i * i + j * j
    v = np.rint(np.hypot(i, j))
    keep = ~((i == 0) & (j < 0)) & (r2 < rU * rU) & (r2 >= rL * rL) & (v < rU) & (v >= rL)
    return set(zip(i[keep].tolist(), j[keep].tolist()))


@pytest.mark.parametrize("N,rU,K", [(64, 24, 871), (128, 50, 3853), (160, 64, 6323),
                                    (256, 100, 15559)])
def test_pixel_set_brute_force(N, rU, K):
    iCol, iRow, iSig, iPxl = oracle.pixel_set(N, rU)
    assert len(iCol) == K
    assert set(zip(iCol.tolist(), iRow.tolist())) == _brute_pixels(N, rU)
    # shells are the rounded radius; iPxl is the half-complex row-major index
    np.testing.assert_array_equal(iSig, np.rint(np.hypot(iCol, iRow)).astype(int))
    np.testing.assert_array_equal(iPxl, np.where(iRow >= 0, iRow, iRow + N) * (N // 2 + 1) + iCol)
    # loop order: j outer ascending, i inner ascending
    key = iRow * 10000 + iCol
    assert np.all(np.diff(key) > 0)


def test_pixel_set_lower_cutoff():
    iCol, iRow, iSig, _ = oracle.pixel_set(64, 20, 5)
    assert set(zip(iCol.tolist(), iRow.tolist())) == _brute_pixels(64, 20, 5)
    assert iSig.min() >= 5 and iSig.max() < 20


def test_workload_pixel_set_equals_oracle():
    for N, rU in [(64, 24), (128, 50)]:
        a = wl.pixel_set(N, rU)
        b = oracle.pixel_set(N, rU)
        for x, y in zip(a, b[:3]):
            np.testing.assert_array_equal(x, y)


# ---------------------------------------------------------------------------
# a2 CTF
# ---------------------------------------------------------------------------

def _lambda_codata(volt):
    h, m0, e, c = 6.62607015e-34, 9.1093837015e-31, 1.602176634e-19, 299792458.0
    lam = h / math.sqrt(2 * m0 * e * volt * (1 + e * volt / (2 * m0 * c * c)))
    return lam * 1e10    # Angstrom


def test_ctf_at_origin_is_amplitude_contrast():
    # u = 0: chi = -phaseShift -> ctf = -sqrt(1-A^2) sin(-ps) + A cos(ps)
    for A, ps in [(0.07, 0.0), (0.1, 0.3)]:
        v = oracle.ctf([0], [0], 128, 1.5, 300e3, 20000, 15000, 0.4, 2.7e7, A, ps)
        assert v[0] == pytest.approx(-math.sqrt(1 - A * A) * math.sin(-ps) + A * math.cos(ps),
                                     abs=1e-6)


def test_ctf_wavelength_matches_codata():
    # Cs = 0, A = 0, no astigmatism: ctf = sin(pi lambda dU u^2); first zero
    # at u^2 = 1 / (lambda dU).  Evaluate on a fine pixel line and locate it.
    N, px, dU = 4096, 0.25, 20000.0
    lam = _lambda_codata(300e3)
    ic = np.arange(1, 400)
    v = oracle.ctf(ic, np.zeros_like(ic), N, px, 300e3, dU, dU, 0.0, 0.0, 0.0)
    u = ic / (N * px)
    chi = np.pi * lam * dU * u * u
    # the reference constant 12.2643247 differs from CODATA by ~6e-6 relative
    np.testing.assert_allclose(v, np.sin(chi), atol=2e-4 + 1e-5 * chi.max())
    assert abs(lam - 0.019687) < 2e-6     # textbook 300 kV wavelength


def test_ctf_astigmatism_symmetry():
    iCol, iRow, _, _ = oracle.pixel_set(128, 50)
    a = oracle.ctf(iCol, iRow, 128, 1.5, 300e3, 25000, 18000, 0.3, 2.7e7, 0.07)
    b = oracle.ctf(iCol, iRow, 128, 1.5, 300e3, 25000, 18000, 0.3 + math.pi, 2.7e7, 0.07)
    np.testing.assert_allclose(a, b, atol=2e-5)
    # swapping dU and dV equals rotating the astigmatism axis by 90 degrees
    c = oracle.ctf(iCol, iRow, 128, 1.5, 300e3, 18000, 25000, 0.3 + math.pi / 2, 2.7e7, 0.07)
    np.testing.assert_allclose(a, c, atol=2e-5)
    # no astigmatism -> depends on |(i, j)| only
    d = oracle.ctf(iCol, iRow, 128, 1.5, 300e3, 20000, 20000, 0.0, 2.7e7, 0.07)
    e = oracle.ctf(iRow * 0 + np.hypot(iCol, iRow).round().astype(int) * 0 + iRow, iCol, 128,
                   1.5, 300e3, 20000, 20000, 0.0, 2.7e7, 0.07)
    sw = np.abs(iCol) == np.abs(iRow)
    np.testing.assert_allclose(d[sw], e[sw], atol=2e-5)


def test_ctf_spherical_aberration_sign():
    # with zero defocus chi = K2 u^4 > 0: ctf = -w1 sin(chi) + A cos(chi)
    lam = _lambda_codata(300e3)
    u = 0.3
    chi = math.pi / 2 * 2.7e7 * lam ** 3 * u ** 4
    N, px = 200, 1.0
    v = oracle.ctf([int(u * N * px)], [0], N, px, 300e3, 0.0, 0.0, 0.0, 2.7e7, 0.07)
    assert v[0] == pytest.approx(-math.sqrt(1 - 0.07 ** 2) * math.sin(chi) + 0.07 * math.cos(chi),
                                 abs=2e-4)


# ---------------------------------------------------------------------------
# a5 translation phases
# ---------------------------------------------------------------------------

def test_translate_properties():
    N = 128
    iCol, iRow, _, _ = oracle.pixel_set(N, 50)
    s1 = oracle.translate(1.5, -2.0, N, iCol, iRow)
    s2 = oracle.translate(-0.5, 3.0, N, iCol, iRow)
    s12 = oracle.translate(1.0, 1.0, N, iCol, iRow)
    z1, z2, z12 = (s[:, 0] + 1j * s[:, 1] for s in (s1, s2, s12))
    np.testing.assert_allclose(np.abs(z1), 1, atol=1e-6)
    np.testing.assert_allclose(z1 * z2, z12, atol=2e-5)
    # shift theorem against numpy's FFT: a delta at (tx, ty) has FT exp(-2 pi i k.t / N)
    img = np.zeros((N, N))
    img[3, 5] = 1.0              # row y = 3, column x = 5
    F = np.fft.rfft2(img)        # F[j, i]
    s = oracle.translate(5.0, 3.0, N, iCol, iRow)
    np.testing.assert_allclose(s[:, 0] + 1j * s[:, 1], F[iRow % N, iCol], atol=1e-5)


# ---------------------------------------------------------------------------
# a4 rotation matrix
# ---------------------------------------------------------------------------

def test_rotmat_axis_angle():
    th = 0.7
    q = [math.cos(th / 2), 0, 0, math.sin(th / 2)]
    M = oracle.rotmat(q)
    c, s = math.cos(th), math.sin(th)
    np.testing.assert_allclose(M, [[c, -s, 0], [s, c, 0], [0, 0, 1]], atol=1e-15)
    q = [math.cos(th / 2), math.sin(th / 2), 0, 0]
    np.testing.assert_allclose(oracle.rotmat(q), [[1, 0, 0], [0, c, -s], [0, s, c]], atol=1e-15)


def test_rotmat_orthonormal_and_composition():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b = rng.normal(size=4), rng.normal(size=4)
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        A, B = oracle.rotmat(a), oracle.rotmat(b)
        np.testing.assert_allclose(A @ A.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(A) == pytest.approx(1.0, abs=1e-14)
        # Hamilton product (quaternion_mul, src/Geometry/Euler.cpp:13-26)
        w = a[0] * b[0] - a[1:] @ b[1:]
        v = a[0] * b[1:] + b[0] * a[1:] + np.cross(a[1:], b[1:])
        np.testing.assert_allclose(oracle.rotmat(np.r_[w, v]), A @ B, atol=1e-14)
        np.testing.assert_allclose(oracle.rotmat(-a), A, atol=1e-15)


# ---------------------------------------------------------------------------
# a6 projection
# ---------------------------------------------------------------------------

def _real_volume(V, seed=0):
    rng = np.random.default_rng(seed)
    vol = rng.normal(size=(V, V, V))
    half = np.fft.rfftn(vol)                    # [z][y][x], x = V/2 + 1
    full = np.fft.fftn(vol)
    hv = np.stack([half.real, half.imag], -1).astype(np.float32)
    return hv, full


def _trilinear_full(full, p):
    """Trilinear interpolation of the full complex spectrum with periodic wrap."""
    V = full.shape[0]
    x0 = np.floor(p).astype(int)
    d = p - x0
    acc = 0j
    for k in range(2):
        for j in range(2):
            for i in range(2):
                w = (d[0] if i else 1 - d[0]) * (d[1] if j else 1 - d[1]) * (d[2] if k else 1 - d[2])
                acc += w * full[(x0[2] + k) % V, (x0[1] + j) % V, (x0[0] + i) % V]
    return acc


def test_project_identity_is_exact_voxels():
    V, N, pf = 32, 16, 2
    hv, full = _real_volume(V)
    iCol, iRow, _, _ = oracle.pixel_set(N, 6)
    out = oracle.project(hv, V, np.eye(3), iCol, iRow, pf)
    ref = hv[0, (iRow * pf) % V, iCol * pf]
    np.testing.assert_array_equal(out, ref)


def test_project_matches_full_spectrum_interpolation():
    V, N, pf = 32, 16, 2
    hv, full = _real_volume(V, 1)
    iCol, iRow, _, _ = oracle.pixel_set(N, 6)
    rng = np.random.default_rng(2)
    for _ in range(6):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        M = oracle.rotmat(q)
        got = oracle.project(hv, V, M, iCol, iRow, pf, f64=True)
        got32 = oracle.project(hv, V, M, iCol, iRow, pf)
        for n in range(len(iCol)):
            p = M @ np.array([iCol[n] * pf, iRow[n] * pf, 0.0])
            ref = _trilinear_full(full.astype(np.complex64).astype(np.complex128), p)
            assert abs(got[n, 0] + 1j * got[n, 1] - ref) <= 1e-9 * (1 + abs(ref))
            assert abs(got32[n, 0] + 1j * got32[n, 1] - ref) <= 1e-4 * (1 + abs(ref))


def test_project_half_turn_conjugates():
    # 180 degrees about y maps (x, y, 0) to (-x, y, 0): the mirrored half-space
    # sample is conj(F(x, -y, 0)).
    V, N, pf = 32, 16, 2
    hv, full = _real_volume(V, 3)
    iCol, iRow, _, _ = oracle.pixel_set(N, 6)
    q = [0.0, 0.0, 1.0, 0.0]
    out = oracle.project(hv, V, oracle.rotmat(q), iCol, iRow, pf)
    ref = hv[0, (-iRow * pf) % V, iCol * pf] * np.array([1, -1], np.float32)
    sel = iCol > 0
    np.testing.assert_allclose(out[sel], ref[sel], atol=1e-5)


# ---------------------------------------------------------------------------
# a7 likelihood and a8 marginals
# ---------------------------------------------------------------------------

def test_logdatavs_zero_at_truth_and_expansion():
    rng = np.random.default_rng(4)
    K, n = 200, 7
    pri = rng.normal(size=(K, 2)).astype(np.float32)
    ctf = rng.uniform(-1, 1, size=(K, n)).astype(np.float32)
    sig = -rng.uniform(0.5, 2, size=(K, n)).astype(np.float32)
    dat = (ctf[..., None] * pri[:, None, :]).astype(np.float32)
    res = oracle.logdatavs_pm(dat, pri, ctf, sig)
    np.testing.assert_allclose(res, 0, atol=1e-6)
    noise = rng.normal(size=(K, n, 2)).astype(np.float32)
    dat2 = (dat + noise).astype(np.float32)
    res2 = oracle.logdatavs_pm(dat2, pri, ctf, sig)
    # expansion: sig |X|^2 + sig c^2 |P|^2 - 2 Re(sig c conj(X) P)
    X = dat2[..., 0] + 1j * dat2[..., 1]
    Pc = pri[:, 0] + 1j * pri[:, 1]
    A = np.sum(sig * np.abs(X) ** 2, 0)
    C = np.sum(sig * ctf ** 2 * np.abs(Pc[:, None]) ** 2, 0)
    B = np.sum(np.real(sig * ctf * np.conj(X) * Pc[:, None]), 0)
    np.testing.assert_allclose(res2, A + C - 2 * B, rtol=1e-5)
    # image-major FP64 form agrees
    for j in range(n):
        v = oracle.logdatavs_f64(dat2[:, j], pri, ctf[:, j], sig[:, j])
        assert v == pytest.approx(res2[j], rel=1e-5)


def test_simd256_likelihood_equals_scalar():
    """SIMD256Float (8 images per AVX vector, deinterleaved pairs, scalar tail)
    gives the scalar pixel-major sum bit for bit, for image counts that are and
    are not multiples of 8 (tail only, vector only, both)."""
    rng = np.random.default_rng(14)
    K = 97
    for n in (3, 8, 16, 21):
        pri = rng.normal(size=(K, 2)).astype(np.float32)
        ctf = rng.uniform(-1, 1, size=(K, n)).astype(np.float32)
        sig = -rng.uniform(0.5, 2, size=(K, n)).astype(np.float32)
        dat = rng.normal(size=(K, n, 2)).astype(np.float32)
        np.testing.assert_array_equal(oracle.logdatavs_pm_simd256(dat, pri, ctf, sig),
                                      oracle.logdatavs_pm(dat, pri, ctf, sig))


@pytest.fixture(scope="module")
def scan_case():
    V, N, pf = 32, 16, 2
    hv, _ = _real_volume(V, 5)
    hv /= np.abs(hv).max()
    iCol, iRow, _, _ = oracle.pixel_set(N, 6)
    K = len(iCol)
    rng = np.random.default_rng(6)
    quat = wl.super_fibonacci(24)
    trans = wl.shift_grid(3, 1.0)
    P = 5
    dat = rng.normal(size=(P, K, 2)).astype(np.float32) * 0.01
    ctf = rng.uniform(-1, 1, size=(P, K)).astype(np.float32)
    sig = np.full((P, K), -2.0, np.float32)
    # plant the true pose (r=7, t=4) into particle 2
    pr = oracle.project(hv, V, oracle.rotmat(quat[7]), iCol, iRow, pf)
    s = oracle.translate(trans[4, 0], trans[4, 1], N, iCol, iRow)
    q = (pr[:, 0] + 1j * pr[:, 1]) * (s[:, 0] + 1j * s[:, 1])
    dat[2, :, 0] += (ctf[2] * q.real).astype(np.float32)
    dat[2, :, 1] += (ctf[2] * q.imag).astype(np.float32)
    pR = rng.uniform(0.5, 1.5, 24)
    pR /= pR.sum()
    pT = rng.uniform(0.5, 1.5, 9)
    pT /= pT.sum()
    return dict(V=V, N=N, pf=pf, hv=hv, iCol=iCol, iRow=iRow, quat=quat, trans=trans, dat=dat,
                ctf=ctf, sig=sig, pR=pR, pT=pT)


def test_global_scan_f64_against_brute_force(scan_case):
    c = scan_case
    wC, wR, wT, base, dvp = oracle.global_scan(c["hv"], c["V"], c["pf"], c["iCol"], c["iRow"],
                                               c["N"], c["quat"], c["trans"], c["dat"],
                                               c["ctf"], c["sig"], c["pR"], c["pT"],
                                               want_dvp=True)
    # brute-force dvp from the oracle's pieces with an independent reduction
    for l in range(len(base)):
        X = c["dat"][l, :, 0] + 1j * c["dat"][l, :, 1]
        for r in (0, 7, 23):
            pr = oracle.project(c["hv"], c["V"], oracle.rotmat(c["quat"][r]), c["iCol"],
                                c["iRow"], c["pf"], f64=True)
            for t in (0, 4, 8):
                ph = np.exp(-2j * np.pi * (c["iCol"] * c["trans"][t, 0] +
                                          c["iRow"] * c["trans"][t, 1]) / c["N"])
                model = c["ctf"][l] * (pr[:, 0] + 1j * pr[:, 1]) * ph
                ref = np.sum(c["sig"][l] * np.abs(X - model) ** 2)
                assert dvp[l, r, t] == pytest.approx(ref, rel=1e-9)
    assert np.unravel_index(dvp[2].argmax(), dvp[2].shape) == (7, 4)
    # marginal identities: sum_r pR wR = wC = sum_t pT wT; max term is 1
    np.testing.assert_allclose(wR @ c["pR"], wC, rtol=1e-12)
    np.testing.assert_allclose(wT @ c["pT"], wC, rtol=1e-12)
    np.testing.assert_allclose(base, dvp.reshape(len(base), -1).max(1))
    w = np.exp(dvp - base[:, None, None])
    np.testing.assert_allclose(wC, np.einsum("lrt,r,t->l", w, c["pR"], c["pT"]), rtol=1e-12)


def test_global_scan_f32_streaming_matches_f64(scan_case):
    c = scan_case
    args = (c["hv"], c["V"], c["pf"], c["iCol"], c["iRow"], c["N"], c["quat"], c["trans"],
            c["dat"], c["ctf"], c["sig"], c["pR"], c["pT"])
    w64 = oracle.global_scan(*args, f64=True)
    for threads in (1, 4):
        w32 = oracle.global_scan(*args, f64=False, threads=threads)
        np.testing.assert_allclose(w32[3], w64[3], rtol=1e-5)
        np.testing.assert_allclose(w32[0], w64[0], rtol=1e-4)
        np.testing.assert_allclose(w32[1], w64[1], rtol=1e-4, atol=1e-6 * w64[1].max())
        np.testing.assert_allclose(w32[2], w64[2], rtol=1e-4, atol=1e-6 * w64[2].max())


def test_global_scan_simd256_equals_scalar(scan_case):
    """The scan with SIMD256Float's sum (the reference's ENABLE_SIMD_256 build)
    equals the scalar scan bit for bit on one thread (same accumulation
    order), on a case with 13 particles (one vector and a tail)."""
    c = scan_case
    rng = np.random.default_rng(2)
    P = 13
    dat = np.concatenate([c["dat"], rng.normal(size=(P - 5,) + c["dat"].shape[1:]) * 0.01])
    ctf = np.concatenate([c["ctf"], rng.uniform(-1, 1, size=(P - 5, c["ctf"].shape[1]))])
    sig = np.full((P, c["ctf"].shape[1]), -2.0)
    args = (c["hv"], c["V"], c["pf"], c["iCol"], c["iRow"], c["N"], c["quat"], c["trans"],
            dat, ctf, sig, c["pR"], c["pT"])
    a = oracle.global_scan(*args, f64=False, threads=1)
    b = oracle.global_scan(*args, f64=False, threads=1, simd=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_local_scan_reduces_to_global(scan_case):
    """One local phase whose per-particle lists are the whole grid equals the
    global scan (pinned above by brute force); with ragged sub-lists it equals
    the marginals of the brute-force dvp restricted to each particle's poses."""
    c = scan_case
    args = (c["hv"], c["V"], c["pf"], c["iCol"], c["iRow"], c["N"])
    wC0, wR0, wT0, base0, dvp = oracle.global_scan(*args, c["quat"], c["trans"], c["dat"],
                                                   c["ctf"], c["sig"], c["pR"], c["pT"],
                                                   want_dvp=True)
    P, R, T = dvp.shape
    full = oracle.local_scan(*args, np.tile(c["quat"], (P, 1)), np.full(P, R),
                             np.tile(c["trans"], (P, 1)), np.full(P, T), np.tile(c["pR"], P),
                             np.tile(c["pT"], P), c["dat"], c["ctf"], c["sig"])
    np.testing.assert_allclose(full[3], base0, rtol=1e-12)
    np.testing.assert_allclose(full[0], wC0, rtol=1e-10)
    np.testing.assert_allclose(full[1].reshape(P, R), wR0, rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(full[2].reshape(P, T), wT0, rtol=1e-10, atol=1e-300)
    rng = np.random.default_rng(9)
    rs = [np.sort(rng.choice(R, rng.integers(1, R + 1), replace=False)) for _ in range(P)]
    ts = [np.sort(rng.choice(T, rng.integers(1, T + 1), replace=False)) for _ in range(P)]
    rs[2], ts[2] = np.array([3, 7, 11]), np.array([0, 4])
    wC, wR, wT, base, bR, bT = oracle.local_scan(
        *args, np.concatenate([c["quat"][r] for r in rs]), [len(r) for r in rs],
        np.concatenate([c["trans"][t] for t in ts]), [len(t) for t in ts],
        np.concatenate([c["pR"][r] for r in rs]), np.concatenate([c["pT"][t] for t in ts]),
        c["dat"], c["ctf"], c["sig"])
    ro = np.r_[0, np.cumsum([len(r) for r in rs])]
    to = np.r_[0, np.cumsum([len(t) for t in ts])]
    for l in range(P):
        sub = dvp[l][np.ix_(rs[l], ts[l])]
        assert base[l] == pytest.approx(sub.max(), rel=1e-12)
        w = np.exp(sub - sub.max())
        np.testing.assert_allclose(wR[ro[l]:ro[l + 1]], w @ c["pT"][ts[l]], rtol=1e-10)
        np.testing.assert_allclose(wT[to[l]:to[l + 1]], c["pR"][rs[l]] @ w, rtol=1e-10)
        assert wC[l] == pytest.approx(c["pR"][rs[l]] @ w @ c["pT"][ts[l]], rel=1e-10)
        assert (bR[l], bT[l]) == np.unravel_index(sub.argmax(), sub.shape)
    assert (rs[2][bR[2]], ts[2][bT[2]]) == (7, 4)


# ---------------------------------------------------------------------------
# a10 systematic resampling
# ---------------------------------------------------------------------------

def test_resample_hand_worked():
    w = np.array([0.1, 0.2, 0.3, 0.4])
    u = np.ones(4)
    idx, wn = oracle.resample(w, u, 4, 0.05)
    np.testing.assert_array_equal(idx, [0, 1, 2, 3])
    idx, wn = oracle.resample(w, u, 4, 0.2)
    np.testing.assert_array_equal(idx, [1, 2, 3, 3])
    np.testing.assert_allclose(wn, 0.25)
    # prior weights multiply the likelihood; new weight is 1/u (normalised)
    idx, wn = oracle.resample(np.ones(4), np.array([1.0, 1.0, 6.0, 2.0]), 5, 0.1)
    np.testing.assert_array_equal(idx, [0, 2, 2, 2, 3])
    np.testing.assert_allclose(wn, np.array([1, 1 / 6, 1 / 6, 1 / 6, 1 / 2]) /
                               (1 + 3 / 6 + 1 / 2))


def test_resample_counts_kitagawa():
    rng = np.random.default_rng(9)
    for _ in range(20):
        n_in, n_out = rng.integers(2, 40), rng.integers(1, 60)
        w = rng.uniform(0, 1, n_in)
        u = rng.uniform(0.1, 1, n_in)
        u0 = rng.uniform(0, 1.0 / n_out)
        idx, _ = oracle.resample(w, u, n_out, u0)
        assert np.all(np.diff(idx) >= 0)
        W = w * u / np.sum(w * u)
        cnt = np.bincount(idx, minlength=n_in)
        assert np.all(cnt >= np.floor(n_out * W) - 1) and np.all(cnt <= np.ceil(n_out * W) + 1)
        assert cnt.sum() == n_out


# ---------------------------------------------------------------------------
# synthetic volume closed form
# ---------------------------------------------------------------------------

def test_blob_ft_matches_numpy_fft():
    N, pf = 32, 2
    V = N * pf
    sigma = 2.0
    centres = np.array([[3.0, -2.0, 5.0], [-4.0, 1.0, 0.0]])
    amps = np.array([1.0, 0.5])
    # real-space blobs on the padded grid with the box centre at index 0
    g = np.fft.fftfreq(V, 1.0 / V)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vol = np.zeros((V, V, V))
    for c, a in zip(centres, amps):
        vol += a * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * sigma ** 2))
    F = np.fft.rfftn(vol)
    rng = np.random.default_rng(3)
    ix = rng.integers(0, V // 2 + 1, 200)
    iy = rng.integers(-V // 2, V // 2, 200)
    iz = rng.integers(-V // 2, V // 2, 200)
    got = oracle.blob_ft(V, N, centres, amps, sigma, ix, iy, iz)
    ref = F[iz % V, iy % V, ix]
    scale = np.abs(F).max()
    np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], ref, atol=1e-5 * scale)


# ---------------------------------------------------------------------------
# a9 / a11 particle update: inferACG, calVari, balanceWeight, peaks, shifts
# ---------------------------------------------------------------------------

def _qmat(m):
    """4x4 matrix of left multiplication by quaternion m (x -> m x)."""
    w, x, y, z = m
    return np.array([[w, -x, -y, -z], [x, w, -z, y], [y, z, w, -x], [z, -y, x, w]])


def _acg_samples(rng, lam, n, m=None):
    v = rng.normal(size=(n, 4)) * np.sqrt(lam)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if m is not None:
        v = v @ _qmat(m).T
    return v


def test_infer_acg_recovers_shape_and_is_equivariant():
    rng = np.random.default_rng(21)
    lam = np.array([1.0, 0.3, 0.1, 0.05])
    q = _acg_samples(rng, lam, 40000)
    A, it = oracle.infer_acg(q)
    # Tyler's fixed point estimates the ACG matrix up to scale (trace 4)
    assert np.trace(A) == pytest.approx(4.0, abs=2e-3)
    np.testing.assert_allclose(np.diag(A) / A[0, 0], lam / lam[0], rtol=0.05, atol=2e-3)
    assert np.abs(A - np.diag(np.diag(A))).max() < 0.02
    # equivariance: samples m q give M A M^T
    m = np.array([0.3, -0.5, 0.7, 0.4])
    m /= np.linalg.norm(m)
    M = _qmat(m)
    A2, _ = oracle.infer_acg(q @ M.T)
    np.testing.assert_allclose(A2, M @ A @ M.T, atol=5e-4)
    # the same frame change through the pre-multiplication argument
    A3, _ = oracle.infer_acg(q, pre=m)
    np.testing.assert_allclose(A3, A2, atol=1e-12)
    # q and -q are the same rotation
    sgn = np.where(rng.uniform(size=(len(q), 1)) < 0.5, -1.0, 1.0)
    A4, _ = oracle.infer_acg(q * sgn)
    np.testing.assert_allclose(A4, A, atol=1e-12)


def test_cal_vari_rot_recovers_mean_and_concentrations():
    rng = np.random.default_rng(22)
    k = np.array([0.02, 0.05, 0.01])
    m = np.array([0.1, 0.8, -0.3, 0.5])
    m /= np.linalg.norm(m)
    q = _acg_samples(rng, np.r_[1.0, k], 30000, m)
    mean, kk = oracle.cal_vari_rot(q)
    assert abs(abs(mean @ m) - 1) < 1e-4
    np.testing.assert_allclose(kk, k, rtol=0.05)
    ev = np.linalg.eigvalsh(oracle.infer_acg(q)[0])
    A, _ = oracle.infer_acg(q)
    v = oracle.sym4_top(A)
    np.testing.assert_allclose(A @ v, ev[-1] * v, atol=1e-10)


def test_balance_rot_makes_samples_uniform():
    rng = np.random.default_rng(23)
    q = _acg_samples(rng, np.array([1.0, 0.2, 0.1, 0.3]), 60000)
    w = oracle.balance_rot(q)
    assert w.sum() == pytest.approx(1.0, rel=1e-12)
    # weights 1/pdf turn ACG samples into uniform ones on S^3: E[q q^T] = I / 4
    S = (q * w[:, None]).T @ q
    np.testing.assert_allclose(S, np.eye(4) / 4, atol=0.02)


def test_trans_stats_against_numpy_scipy():
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(24)
    t = rng.normal([0.5, -1.0], [2.0, 0.7], size=(53, 2))
    m, s, w = oracle.trans_stats(t)
    np.testing.assert_allclose(m, t.mean(0), rtol=1e-12)
    np.testing.assert_allclose(s, t.std(0, ddof=1), rtol=1e-12)
    pdf = multivariate_normal(m, np.diag(s ** 2)).pdf(t)
    np.testing.assert_allclose(w, (1 / pdf) / np.sum(1 / pdf), rtol=1e-10)


def test_keep_peak_hand_worked():
    u = np.array([0.1, 4.0, 2.0, 1.0, 3.0, 0.5, 4.0, 0.2, 1.5, 0.05, 2.5, 0.3, 0.7, 1.2, 0.9, 0.8])
    # 16 samples: rank 16 / 8 = 2 of the descending order is 3.0, so the
    # factor is clamp(3.0 / 4.0) = 0.5 and the half height is 2.0
    out, f, top = oracle.keep_peak(u, set_factor=True)
    assert f == 0.5 and top == 1
    np.testing.assert_allclose(out, np.where(u < 2.0, 0.0, u - 2.0))
    out, f, _ = oracle.keep_peak(u, factor=1e-3)
    assert f == 1e-3
    np.testing.assert_allclose(out, np.where(u < 4e-3, 0.0, u - 4e-3))
    # a flat set clamps at 0.5; a single spike at 1e-3
    assert oracle.keep_peak(np.ones(40), set_factor=True)[1] == 0.5
    assert oracle.keep_peak(np.r_[1.0, np.zeros(39)], set_factor=True)[1] == 1e-3


# ---- projectee preparation (Projector::setProjectee + gridCorrection) ----

def _tik(r, pf, V):
    a = math.pi * r / (pf * V)
    return 1.0 if a == 0 else (math.sin(a) / a) ** 2


def test_projectee_flat_spectrum_stays_flat():
    """A flat half spectrum is a unit delta at the origin after the scaled
    inverse transform; the origin is not moved by padding and its correction
    is j0(0)^2 = 1, so the padded spectrum is flat too (pins the 1/N^3 scale
    and the origin of the padding)."""
    N, pf = 8, 2
    out = oracle.projectee(np.ones((N, N, N // 2 + 1)), N, pf)
    assert out.shape == (16, 16, 9)
    np.testing.assert_allclose(out, 1.0, atol=1e-12)


@pytest.mark.parametrize("x0", [(2, -3, 1), (-4, 3, -2)])
def test_projectee_shifted_delta_closed_form(x0):
    """A unit delta at signed voxel x0 = (i, j, k): input spectrum
    exp(-2 pi i k.x0 / N), output exp(-2 pi i k'.x0 / V) / j0(pi |x0| / (pf V))^2.
    Pins the transform sign, the wrap of negative indices at both sizes and
    the correction's argument r / (pf V) with V the padded size."""
    N, pf = 8, 2
    V = pf * N
    i0, j0_, k0 = x0
    kz, ky, kx = np.meshgrid(np.arange(N), np.arange(N), np.arange(N // 2 + 1), indexing="ij")
    ft = np.exp(-2j * np.pi * (kx * i0 + ky * j0_ + kz * k0) / N)
    out = oracle.projectee(ft, N, pf)
    Kz, Ky, Kx = np.meshgrid(np.arange(V), np.arange(V), np.arange(V // 2 + 1), indexing="ij")
    ref = np.exp(-2j * np.pi * (Kx * i0 + Ky * j0_ + Kz * k0) / V)
    ref /= _tik(math.sqrt(i0 * i0 + j0_ * j0_ + k0 * k0), pf, V)
    np.testing.assert_allclose(out, ref, atol=1e-12)


def test_projectee_brute_force_dft():
    """N = 4, pf = 2 on an asymmetric random map: every step by explicit sums
    (DFT definitions, padding loop, TIK_RL), independent of numpy's FFT
    conventions; pins the axis order (z, y, x) and the half dimension (x)."""
    N, pf = 4, 2
    V = pf * N
    rng = np.random.default_rng(17)
    f = rng.normal(size=(N, N, N))               # f[z][y][x]
    n = np.arange(N)
    E = np.exp(-2j * np.pi * np.outer(n, n) / N)
    ft = np.einsum("zyx,az,by,cx->abc", f, E, E, E)[:, :, :N // 2 + 1]
    pad = np.zeros((V, V, V))
    for z in range(N):
        for y in range(N):
            for x in range(N):
                k, j, i = (z if z < N // 2 else z - N), (y if y < N // 2 else y - N), \
                    (x if x < N // 2 else x - N)
                r = math.sqrt(i * i + j * j + k * k)
                pad[k % V, j % V, i % V] = f[z, y, x] / _tik(r, pf, V)
    m = np.arange(V)
    EV = np.exp(-2j * np.pi * np.outer(m, m) / V)
    ref = np.einsum("zyx,az,by,cx->abc", pad, EV, EV, EV)[:, :, :V // 2 + 1]
    np.testing.assert_allclose(oracle.projectee(ft, N, pf), ref, atol=1e-10)


# ---- particle images (Optimiser::initImg image path + allocPreCal datP) ----

def _bg_image(N, r):
    """An image whose background (outside r) alternates +1 / -1 in raster
    order, so its mean and sample standard deviation are known in closed
    form, and whose inside is zero."""
    h = N // 2
    img = np.zeros((N, N))
    sgn, n = 1.0, 0
    for y in range(N):
        for x in range(N):
            if (x - h) ** 2 + (y - h) ** 2 > r * r:
                img[y, x] = sgn
                sgn = -sgn
                n += 1
    return img, n


def test_particle_images_brute_force_dft():
    """N = 8: the normalisation with hand-computed background statistics, the
    centre-to-origin move by an explicit index loop and the DFT by explicit
    sums; pins the roll direction, the transform sign, the missing
    normalisation and the wrap of negative rows in the packing."""
    N, r = 8, 3.0
    h = N // 2
    rng = np.random.default_rng(3)
    img, n = _bg_image(N, r)
    inner = rng.normal(size=(N, N))
    yy, xx = np.mgrid[0:N, 0:N]
    inside = (xx - h) ** 2 + (yy - h) ** 2 <= r * r
    img[inside] = inner[inside]
    bgv = img[~inside]
    mean = bgv.sum() / n
    sd = math.sqrt(((bgv - mean) ** 2).sum() / (n - 1))
    stored = np.zeros((N, N))
    for y in range(N):
        for x in range(N):
            stored[(y + h) % N, (x + h) % N] = (img[y, x] - mean) / sd
    iCol = np.array([0, 1, 3, 4, 2, 0])
    iRow = np.array([0, -1, 2, -4, 3, -2])
    ref = np.zeros(len(iCol), complex)
    for k in range(len(iCol)):
        kx, ky = iCol[k], iRow[k]
        for y in range(N):
            for x in range(N):
                ref[k] += stored[y, x] * np.exp(-2j * np.pi * (kx * x + ky * y) / N)
    out = oracle.particle_images(img[None], r, iCol, iRow)[0]
    np.testing.assert_allclose(out, ref, atol=1e-10)


@pytest.mark.parametrize("a,b", [(2, -1), (-3, 2)])
def test_particle_images_shifted_delta(a, b):
    """A delta of height A at centred pixel (a, b) over the +1 / -1
    background: after normalisation the stored image is the delta at (a, b)
    (wrapped) plus the known background, so every coefficient is
    ((A - m) / sd) exp(-2 pi i (kx a + ky b) / N) plus the background's own
    transform; the background part is cancelled by subtracting the same
    image without the delta."""
    N, r, A = 16, 5.0, 7.5
    h = N // 2
    img, n = _bg_image(N, r)
    bgv = img[img != 0]
    m = bgv.mean()
    sd = bgv.std(ddof=1)
    with_delta = img.copy()
    with_delta[h + b, h + a] = A
    iCol, iRow = oracle.pixel_set(N, 6)[:2]
    d = oracle.particle_images(np.stack([with_delta, img]), r, iCol, iRow)
    diff = d[0] - d[1]
    ref = (A - 0.0) / sd * np.exp(-2j * np.pi * (iCol * a + iRow * b) / N)
    np.testing.assert_allclose(diff, ref, atol=1e-10)
    assert abs(m) < 1 and n > 2


def test_init_sigma_explicit_loops():
    """N = 8, three images: initSigma written out with explicit loops
    (normalisation, centre move, DFT sums over the stored half spectrum,
    powerSpectrum with its radius test and counters, ringAverage of Re + Im
    with its own counter) against the vectorised oracle."""
    N, rbg, P = 8, 3.0, 3
    h, r = N // 2, N // 2 - 1
    rng = np.random.default_rng(8)
    imgs = rng.normal(1.0, 2.0, size=(P, N, N))
    spectra = []
    for p in range(P):
        bgv = [imgs[p, y, x] for y in range(N) for x in range(N)
               if (x - h) ** 2 + (y - h) ** 2 > rbg * rbg]
        mean = sum(bgv) / len(bgv)
        sd = math.sqrt(sum((v - mean) ** 2 for v in bgv) / (len(bgv) - 1))
        stored = np.zeros((N, N))
        for y in range(N):
            for x in range(N):
                stored[(y + h) % N, (x + h) % N] = (imgs[p, y, x] - mean) / sd
        F = {}
        for j in range(-h, h):
            for i in range(h + 1):
                F[(i, j)] = sum(stored[y, x] * np.exp(-2j * np.pi * (i * x + j * y) / N)
                                for y in range(N) for x in range(N))
        spectra.append(F)
    avg_ps = np.zeros(r)
    for F in spectra:
        ps, cnt = np.zeros(r), np.zeros(r)
        for (i, j), f in F.items():
            if i * i + j * j < r * r:
                u = int(round(math.hypot(i, j)))
                if u < r:
                    ps[u] += abs(f) ** 2
                    cnt[u] += 1
        avg_ps += ps / cnt / P
    avg = {key: sum(F[key] for F in spectra) / P for key in spectra[0]}
    ps_avg = np.zeros(r)
    for u in range(r):
        acc, c = 0.0, 0
        for (i, j), f in avg.items():
            if int(round(math.hypot(i, j))) == u:
                acc += f.real + f.imag
                c += 1
        ps_avg[u] = (acc / c) ** 2
    sig_ref = (avg_ps - ps_avg) / 2
    sig, rcp = oracle.init_sigma(imgs, rbg, r)
    np.testing.assert_allclose(sig, sig_ref, rtol=1e-10)
    np.testing.assert_allclose(rcp, -0.5 / sig_ref, rtol=1e-10)


def test_init_sigma_white_noise():
    """Images of white Gaussian noise (any mean and spread): after the
    background normalisation every pixel has unit variance, so each
    unnormalised coefficient has E|F|^2 = N^2 and the average image tends to
    zero: sigma^2(u) = N^2 / 2 for every shell u >= 1 (shell 0 is the DC term,
    reduced by the mean subtraction)."""
    N, P = 32, 300
    rng = np.random.default_rng(32)
    imgs = rng.normal(5.0, 2.0, size=(P, N, N))
    sig, rcp = oracle.init_sigma(imgs, 10.0, N // 2 - 1)
    j = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    shell = np.rint(np.hypot(np.arange(N // 2 + 1)[None, :], j)).astype(int)
    cnt = np.array([(shell == u).sum() for u in range(N // 2 - 1)])
    rel = sig[1:] / (N * N / 2) - 1
    # relative spread of a mean of P * cnt / 2 independent exponential powers
    assert np.all(np.abs(rel) < 6 / np.sqrt(P * cnt[1:] / 2)), rel
    pooled = (sig[1:] * cnt[1:]).sum() / cnt[1:].sum() / (N * N / 2)
    assert abs(pooled - 1) < 6 / np.sqrt(P * cnt[1:].sum() / 2), pooled
    np.testing.assert_allclose(rcp, -0.5 / sig)


def test_soft_mask_profile():
    """The mask at hand-worked radii: 1 inside r (and at u = r), 0.5 halfway
    along the 6-pixel edge, 0 at and beyond r + 6, (1 + cos(pi / 3)) / 2 =
    0.75 a third of the way; radially symmetric about pixel (N/2, N/2)."""
    N, r = 32, 8.0
    m = oracle.soft_mask(N, r)
    h = N // 2
    assert m[h, h] == 1 and m[h, h + 8] == 1 and m[h + 7, h] == 1
    assert m[h, h + 11] == pytest.approx(0.5, abs=1e-15)
    assert m[h + 10, h] == pytest.approx(0.75, abs=1e-15)
    assert m[h, h + 14] == pytest.approx(0.0, abs=1e-15) and m[h, h + 15] == 0 and m[0, 0] == 0
    assert m[h - 11, h] == m[h, h - 11] == m[h + 11, h]


def test_masked_images_explicit_loops():
    """N = 12, radius 2 (so the cosine edge and the zero region both fall
    inside the image): the masked image written pixel by pixel with math.cos
    and its DFT by explicit sums at packed coefficients, against the oracle."""
    N, r = 12, 2.0
    h = N // 2
    rng = np.random.default_rng(12)
    img = rng.normal(0.5, 1.5, size=(N, N))
    bgv = [img[y, x] for y in range(N) for x in range(N) if (x - h) ** 2 + (y - h) ** 2 > r * r]
    mean = sum(bgv) / len(bgv)
    sd = math.sqrt(sum((v - mean) ** 2 for v in bgv) / (len(bgv) - 1))
    stored = np.zeros((N, N))
    for y in range(N):
        for x in range(N):
            u = math.hypot(x - h, y - h)
            w = 1.0 if u < r else (0.0 if u > r + 6 else 0.5 + 0.5 * math.cos((u - r) / 6 * math.pi))
            stored[(y + h) % N, (x + h) % N] = (img[y, x] - mean) / sd * w
    iCol = np.array([0, 1, 2, 5, 6, 3])
    iRow = np.array([0, -1, 4, -6, 2, -3])
    ref = np.array([sum(stored[y, x] * np.exp(-2j * np.pi * (iCol[k] * x + iRow[k] * y) / N)
                        for y in range(N) for x in range(N)) for k in range(len(iCol))])
    out = oracle.particle_images(img[None], r, iCol, iRow, masked=True)[0]
    np.testing.assert_allclose(out, ref, atol=1e-10)


# ---------------------------------------------------------------------------
# a9: the post-scan composite, hand-worked
# ---------------------------------------------------------------------------

def _post_scan_case():
    rng = np.random.default_rng(12)
    gq = rng.normal(size=(8, 4))
    gq /= np.linalg.norm(gq, axis=1, keepdims=True)
    gt = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 2.0], [-1.0, -1.0], [3.0, 1.0]])
    # class 1 carries the marginals worked below; classes 0 and 2 other rows
    wR = np.array([[1, 1, 1, 1, 1, 1, 9, 1], [6, 3, 10, 2, 9, 7, 0, 8], [9, 0, 0, 0, 0, 0, 0, 1.0]])
    wT = np.array([[0, 0, 0, 5, 0], [1, 4, 2, 0, 3], [0, 5, 0, 0, 0.0]])
    wC = np.array([2.7, 1.0, 0.5])
    base = np.array([-10.0, -9.0, -12.0])
    return gq, gt, wR, wT, wC, base


def test_post_scan_hand_worked():
    """Classes: u = wC exp(base - max base) = (2.7/e, 1, 0.5/e^3); half-height
    peak at 0.99 leaves (0.00327, 0.01, 0), cdf (0.2465, 1, 1); the draws
    u0C + j/3 = 0.05, 0.383, 0.717 give classes (0, 1, 1), so pick 0.1 -> class 0
    and pick 0.5 -> class 1.  Rotations of class 1: (6 3 10 2 9 7 0 8), peak
    factor clamp(9/10) = 0.5, kept (1 0 5 0 4 2 0 3), cdf / 15 = (1 1 6 6 10 12
    12 15) / 15; u0R = 0.06 over 6 draws -> indices (0 2 2 4 5 7), weights
    1/u normalised = (60 12 12 15 30 20) / 149.  Shifts of class 1 (1 4 2 0 3),
    no peak keeping, cdf (.1 .5 .7 .7 1); u0T = 0.2/3 over 3 draws -> (0 1 4),
    weights (12 3 4) / 19; samples (0,0), (1,0), (3,1): deviations sqrt(7/3) and
    sqrt(1/3), the second raised to minStdT / pfS = 0.6."""
    gq, gt, wR, wT, wC, base = _post_scan_case()
    pR, pT = np.full(8, 1 / 8), np.full(5, 1 / 5)
    kw = dict(min_std_r=0.0, min_std_t=0.9, pfS=1.5)
    a = oracle.post_scan(wC, base, wR, wT, pR, pT, gq, gt, 6, 3, 0.05, 0.1, 0.06, 0.2 / 3, **kw)
    assert a["cls"] == 0
    b = oracle.post_scan(wC, base, wR, wT, pR, pT, gq, gt, 6, 3, 0.05, 0.5, 0.06, 0.2 / 3, **kw)
    assert b["cls"] == 1
    np.testing.assert_array_equal(b["idxR"], [0, 2, 2, 4, 5, 7])
    np.testing.assert_allclose(b["wR"], np.array([60, 12, 12, 15, 30, 20]) / 149, rtol=1e-14)
    np.testing.assert_array_equal(b["q"], gq[[0, 2, 2, 4, 5, 7]])
    assert b["peakR"] == 0.5
    np.testing.assert_array_equal(b["idxT"], [0, 1, 4])
    np.testing.assert_allclose(b["wT"], np.array([12, 3, 4]) / 19, rtol=1e-14)
    np.testing.assert_array_equal(b["t"], gt[[0, 1, 4]])
    np.testing.assert_allclose(b["s"], [np.sqrt(7 / 3), 0.6], rtol=1e-14)
    # rotation concentrations: calVari of the drawn samples (pinned on its
    # own), each raised to (minStdR / pfS)^2; a bound between the smallest and
    # largest k clamps exactly the ones below it
    _, k_raw = oracle.cal_vari_rot(gq[[0, 2, 2, 4, 5, 7]], max_iter=1000)
    kmin = 0.5 * (k_raw.min() + k_raw.max())
    c = oracle.post_scan(wC, base, wR, wT, pR, pT, gq, gt, 6, 3, 0.05, 0.5, 0.06, 0.2 / 3,
                         min_std_r=1.5 * np.sqrt(kmin), min_std_t=0.9, pfS=1.5)
    np.testing.assert_allclose(c["k"], np.maximum(k_raw, kmin), rtol=1e-12)
    assert np.any(np.isclose(c["k"], kmin, rtol=1e-12)) and np.any(c["k"] > kmin)


def test_post_scan_single_class_ignores_class_draws():
    gq, gt, wR, wT, wC, base = _post_scan_case()
    pR, pT = np.full(8, 1 / 8), np.full(5, 1 / 5)
    one = oracle.post_scan(wC[1:2], base[1:2], wR[1:2], wT[1:2], pR, pT, gq, gt, 6, 3, 0.3, 0.9,
                           0.06, 0.2 / 3, 0.0, 0.9, 1.5)
    assert one["cls"] == 0
    np.testing.assert_array_equal(one["idxR"], [0, 2, 2, 4, 5, 7])
    np.testing.assert_array_equal(one["idxT"], [0, 1, 4])


# ---------------------------------------------------------------------------
# a13: the convergence rule, hand-worked
# ---------------------------------------------------------------------------

def _seq(vR, vT, n):
    """k with (k1 k2 k3)^(1/6) = vR (all equal: k = vR^2), s0 = s1 = sqrt(vT)."""
    vR = np.broadcast_to(np.asarray(vR, float), (n,))
    vT = np.broadcast_to(np.asarray(vT, float), (n,))
    return np.repeat((vR ** 2)[:, None], 3, 1), np.repeat(np.sqrt(vT)[:, None], 2, 1)


def test_converge_hand_worked():
    # phases 1..30 after a global scan; checks start at phase 10 (index 9)
    n = 30
    # constant: the first check passes (bests at DBL_MAX), the next fails -> 11
    assert oracle.converge(*_seq(0.1, 0.5, n)) == 11
    # vR falls by 10 % per phase up to phase 15, then stays: stops at 16
    vR = 0.9 ** np.minimum(np.arange(1, n + 1), 15)
    assert oracle.converge(*_seq(vR, 0.5, n)) == 16
    # a 4 % fall per phase is not below 0.95 of the best: stops at 11
    assert oracle.converge(*_seq(0.96 ** np.arange(1, n + 1), 0.5, n)) == 11
    # vR constant, vT falling 10 % per phase up to phase 20: stops at 21
    vT = 0.9 ** np.minimum(np.arange(1, n + 1), 20)
    assert oracle.converge(*_seq(0.1, vT, n)) == 21
    # local search: phases 0.., checks from phase 3: stops at 4
    assert oracle.converge(*_seq(0.1, 0.5, n), phase0=0, min_phase=3) == 4
    # comparison with the best so far, not the previous phase: phase 11 rises
    # in vR but vT falls (continue), phase 12 has vR 0.97 of the best (1.0)
    # though 0.65 of the previous: stops at 12
    vR = np.r_[np.ones(9), 1.0, 1.5, 0.97, np.ones(n - 12)]
    vT = np.r_[np.ones(10), 0.5, 0.5, np.full(n - 12, 0.5)]
    assert oracle.converge(*_seq(vR, vT, n)) == 12
    # rho enters variT: s0 = s1 = 1 with rho 0 then 0.6 (variT 1 -> 0.8)
    k = np.full((n, 3), 0.01)
    s = np.zeros((n, 3))
    s[:, :2] = 1.0
    s[10:, 2] = 0.6
    assert oracle.converge(k, s) == 12
    # the defocus deviation counts too
    assert oracle.converge(*_seq(0.1, 0.5, n), sD=0.9 ** np.arange(n)) == -1
    # never past the sequence: -1
    assert oracle.converge(*_seq(0.9 ** np.arange(1, 13), 0.5, 12)) == -1


def test_stat_img_hand_worked():
    """statImg's bgStddev(0, image, r): the background pixels' root mean
    square about 0 with n - 1 (GSL's sd_m), averaged over images.  N = 4,
    r = 1.2: the background is the pixels with centred i^2 + j^2 > 1.44, i.e.
    all but the five of the centre cross (0,0), (+-1,0), (0,+-1): 11 pixels."""
    N = 4
    h = N // 2
    j, i = np.meshgrid(np.arange(N) - h, np.arange(N) - h, indexing="ij")
    bg = (i * i + j * j) > 1.2 ** 2
    assert bg.sum() == 11
    a = np.zeros((N, N))
    a[bg] = 2.0                      # every background pixel 2: rms = 2 sqrt(11 / 10)
    b = np.zeros((N, N))
    b[bg] = np.r_[np.ones(10), 0.0]  # ten ones and a zero: sqrt(10 / 10) = 1
    got = oracle.stat_img(np.stack([a, b]), bg)
    assert got == pytest.approx(0.5 * (2 * np.sqrt(11 / 10) + 1.0), rel=1e-15)


def test_stat_img_after_background_subtraction_is_one():
    """After substractBgImg with the same region, every image's background
    has mean 0 and sample deviation 1, so statImg's deviation about 0 is 1
    and normaliseImg's scale is 1 (up to rounding)."""
    rng = np.random.default_rng(4)
    imgs = rng.normal(rng.uniform(-3, 3), rng.uniform(0.2, 4), size=(6, 32, 32))
    y, bg = oracle.subtract_bg(imgs, 12.0)
    assert oracle.stat_img(y, bg) == pytest.approx(1.0, abs=1e-12)


# ---------------------------------------------------------------------------
# F4: sigma refresh from the best poses (allReduceSigma)
# ---------------------------------------------------------------------------

def _sigma_case(P=60, N=32, seed=0):
    """A blob volume, random poses, shifts and CTF attributes, and the model
    spectra built from the separately pinned projection, translation and CTF
    functions (not from refresh_sigma's own formula)."""
    rng = np.random.default_rng(seed)
    pf, V = 2, 2 * N
    ix, iy, iz = np.meshgrid(np.arange(V // 2 + 1), np.arange(V), np.arange(V), indexing="ij")
    ix, iy, iz = ix.ravel(), np.where(iy.ravel() < V // 2, iy.ravel(), iy.ravel() - V), \
        np.where(iz.ravel() < V // 2, iz.ravel(), iz.ravel() - V)
    cen = rng.uniform(-6, 6, size=(12, 3))
    ft = oracle.blob_ft(V, N, cen, np.ones(12), 1.5, ix, iy, iz)
    vol = np.zeros((V, V, V // 2 + 1, 2), np.float32)
    vol[iz % V, iy % V, ix] = ft
    quat = rng.normal(size=(P, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    trans = rng.uniform(-2, 2, size=(P, 2))
    attr = np.stack([np.full(P, 300e3), rng.uniform(1e4, 3e4, P), rng.uniform(1e4, 3e4, P),
                     rng.uniform(0, 3, P), np.full(P, 2.7e7), np.full(P, 0.07), np.zeros(P)], 1)
    return rng, vol, V, pf, quat, trans, attr


def _model_spectra(vol, V, pf, N, quat, trans, attr, r, px):
    i, j, row = oracle.spectrum_pixels(N)
    inr = i * i + j * j < r * r
    out = np.zeros((len(quat), N, N // 2 + 1), np.complex128)
    for l in range(len(quat)):
        a = attr[l].astype(np.float32)
        c = oracle.ctf(i[inr], j[inr], N, px, *a)
        s = oracle.translate(trans[l, 0], trans[l, 1], N, i[inr], j[inr])
        p = oracle.project(vol, V, oracle.rotmat(quat[l]), i[inr], j[inr], pf, f64=True)
        m = c * (s[:, 0] + 1j * s[:, 1]) * (p[:, 0] + 1j * p[:, 1])
        out[l][row[inr], i[inr]] = m
    return out


def test_refresh_sigma_noise_free_model_gives_zero():
    """Data equal to the model (CTF . shift . slice built from the pinned
    pieces): every residual is zero, the model-to-data power ratio is 1, so
    sigma is 0 in every shell inside the projection radius."""
    N, r, px = 32, 12, 1.5
    rng, vol, V, pf, quat, trans, attr = _sigma_case(P=8)
    spec = _model_spectra(vol, V, pf, N, quat, trans, attr, r, px)
    noisy = spec + rng.normal(size=spec.shape) + 1j * rng.normal(size=spec.shape)
    sig, _ = oracle.refresh_sigma(vol, V, pf, spec, noisy, attr, px, np.zeros(8), quat, trans,
                                  None, None, 1, r, r, 40.0)
    assert np.abs(sig).max() <= 1e-12 * np.abs(spec).max() ** 2


def test_refresh_sigma_without_model_is_alpha_times_half_power():
    """A zero volume: the model is 0, the ratio 0, and sigma = alpha sigN with
    sigN = |c|^2 / 2 for constant data c, alpha = sqrt(pi) maskRadius /
    (N pixelSize) = sqrt(pi) 5 / 64."""
    N, px = 32, 2.0
    _, vol, V, pf, quat, trans, attr = _sigma_case(P=4)
    c = 3.0 - 4.0j
    spec = np.full((4, N, N // 2 + 1), c)
    sig, rcp = oracle.refresh_sigma(np.zeros_like(vol), V, pf, spec, spec, attr, px, np.zeros(4),
                                    quat, trans, None, None, 1, 10, 15, 5.0)
    alpha = np.sqrt(np.pi) * 5.0 / (N * px)
    np.testing.assert_allclose(sig, alpha * 25.0 / 2, rtol=1e-12)
    np.testing.assert_allclose(rcp, -0.5 / (alpha * 12.5), rtol=1e-12)


def test_refresh_sigma_white_noise_statistics():
    """Model plus complex white noise of per-component deviation s: the
    residual at the true pose is the noise, so sigM = sigN = s^2 per shell;
    with alpha = 1 (maskRadius = N pixelSize / sqrt(pi)) sigma = s^2 whatever
    the ratio.  Each shell within 6 standard errors, and two groups with
    different noise levels stay apart."""
    N, r, px, s = 32, 12, 1.5, 0.7
    P = 60
    rng, vol, V, pf, quat, trans, attr = _sigma_case(P=P, seed=3)
    spec = _model_spectra(vol, V, pf, N, quat, trans, attr, r, px)
    grp = (np.arange(P) % 2).astype(np.int32)
    sd = np.where(grp == 0, s, 2 * s)[:, None, None]
    data = spec + sd * (rng.normal(size=spec.shape) + 1j * rng.normal(size=spec.shape))
    n_shell = N // 2 - 1
    sig, _ = oracle.refresh_sigma(vol, V, pf, data, data, attr, px, np.zeros(P), quat, trans,
                                  np.zeros((P, 2)), grp, 2, r, n_shell, N * px / np.sqrt(np.pi))
    i, j, _ = oracle.spectrum_pixels(N)
    u = np.rint(np.hypot(i, j)).astype(int)
    cnt = np.bincount(u[u < n_shell], minlength=n_shell)
    for g, sg in ((0, s), (1, 2 * s)):
        se = sg ** 2 / np.sqrt(cnt * P / 2)
        assert np.all(np.abs(sig[g] - sg ** 2) <= 6 * se), (g, sig[g] / sg ** 2)
    # a pose with the wrong shift sign leaves signal in the low shells
    bad, _ = oracle.refresh_sigma(vol, V, pf, data, data, attr, px, np.zeros(P), quat, -trans,
                                  None, grp, 2, r, n_shell, N * px / np.sqrt(np.pi))
    assert np.all(bad[:, 2:6] > sig[:, 2:6] * 1.5)


# ---------------------------------------------------------------------------
# F2: reconstruction insert (Reconstructor::insertP, trilinear)
# ---------------------------------------------------------------------------

def test_insert_is_the_adjoint_of_projection():
    """Inserting values x_k at the slice points of a rotation is the adjoint of
    taking the slice (the separately pinned trilinear projection p = P V):
    Re <F, V> = Re sum_k conj(x_k) p_k for any volume V, mirrored points
    included; and T's entries sum to sum_k ctf_k^2 w (the trilinear weights of
    a point sum to 1)."""
    rng = np.random.default_rng(1)
    N, pf = 32, 2
    V = N * pf
    iCol, iRow, _, _ = oracle.pixel_set(N, 14)
    K = len(iCol)
    vol = rng.normal(size=(V, V, V // 2 + 1, 2)).astype(np.float32)
    volc = vol[..., 0].astype(np.float64) + 1j * vol[..., 1]
    for trial in range(3):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        M = oracle.rotmat(q)
        x = (rng.normal(size=K) + 1j * rng.normal(size=K)).astype(np.complex64)
        ctf = np.ones(K, np.float32)
        F = np.zeros((V, V, V // 2 + 1), np.complex128)
        T = np.zeros((V, V, V // 2 + 1))
        oracle.insert(F, T, V, pf, x, ctf, M, iCol, iRow, 1.0)
        p = oracle.project(vol, V, M, iCol, iRow, pf, f64=True)
        p = p[:, 0] + 1j * p[:, 1]
        lhs = np.real(np.sum(np.conj(F) * volc))
        rhs = np.real(np.sum(np.conj(x.astype(np.complex128)) * p))
        assert lhs == pytest.approx(rhs, rel=1e-6)
        c2 = rng.uniform(0.1, 1, K).astype(np.float32)
        T2 = np.zeros_like(T)
        oracle.insert(np.zeros_like(F), T2, V, pf, x, c2, M, iCol, iRow, 0.25)
        assert T2.sum() == pytest.approx(0.25 * np.sum(c2.astype(np.float64) ** 2), rel=1e-12)


def test_insert_on_grid_nodes_and_mirror():
    """The identity rotation puts pixel (i, j) on the padded grid node
    (2i, 2j, 0): all of x there; a half turn about z sends it to (-2i, -2j, 0),
    stored mirrored as conj(x) at (2i, 2j, 0) -- except on the x = 0 plane,
    where -0 is not negative and the point stays at (0, -2j, 0) as x."""
    N, pf = 16, 2
    V = N * pf
    iCol, iRow = np.array([3, 0, 5]), np.array([-2, 4, 1])
    x = np.array([1 + 2j, -3 + 0.5j, 0.25 - 1j], np.complex64)
    turn = np.diag([-1.0, -1.0, 1.0])
    for M in (np.eye(3), turn):
        F = np.zeros((V, V, V // 2 + 1), np.complex128)
        T = np.zeros((V, V, V // 2 + 1))
        oracle.insert(F, T, V, pf, x, np.full(3, 2.0, np.float32), M, iCol, iRow, 1.0)
        for k in range(3):
            if M is turn and iCol[k] == 0:
                row, want = (-2 * iRow[k]) % V, x[k]
            else:
                row, want = (2 * iRow[k]) % V, x[k] if M is not turn else np.conj(x[k])
            assert F[0, row, 2 * iCol[k]] == pytest.approx(2 * want)
            assert T[0, row, 2 * iCol[k]] == pytest.approx(4.0)
        assert np.count_nonzero(F) == 3


# ---------------------------------------------------------------------------
# F4: intensity scale and noise normalisation from best poses
# ---------------------------------------------------------------------------

def test_gsl_median_hand_worked():
    """GSL's 0.5 quantile of sorted data: the middle value for odd n, the mean
    of the two middle values for even n, whatever the input order; rounded
    to float."""
    assert oracle.gsl_median([5.0, 1.0, 3.0]) == 3.0
    assert oracle.gsl_median([4.0, 1.0, 3.0, 2.0]) == 2.5
    assert oracle.gsl_median([7.0]) == 7.0
    assert oracle.gsl_median([0.1, 0.2]) == float(np.float32(0.15000000000000002))


def test_refresh_scale_zero_frequency_hand_worked():
    """Without coordinates only the zero frequency counts, where every
    rotation samples the volume's origin voxel: with M = 3 + 4i there,
    data X = 1 + 2i and the CTF at u = 0 equal to the amplitude contrast A
    (phase shift 0), the plain products give Re(X M) = 3 - 8 = -5 and
    Re(M M) = 9 - 16 = -7, so the scale is (-5 A) / (-7 A^2) = 5 / (7 A).
    With two groups whose data differ by the factor 2 the two scales are
    5 / (7 A) and 10 / (7 A) (median 7.5 / (7 A), no clamp); without groups
    both are the pooled ratio."""
    N, A = 16, 0.5
    V, pf = 2 * N, 2
    vol = np.zeros((V, V, V // 2 + 1, 2), np.float32)
    vol[0, 0, 0] = (3.0, 4.0)
    rng = np.random.default_rng(1)
    P = 4
    spec = rng.normal(size=(P, N, N // 2 + 1)) * 50 + 0j
    spec[:, 0, 0] = [1 + 2j, 2 + 4j, 1 + 2j, 2 + 4j]
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    attr = np.tile([300e3, 2e4, 1.5e4, 0.3, 2.7e7, A, 0.0], (P, 1))
    grp = np.array([0, 1, 0, 1])
    sc = oracle.refresh_scale(vol, V, pf, spec, attr, 1.5, None, q, None, None, grp, 2, 6, 6,
                              0.0, coord=False)
    np.testing.assert_allclose(sc, [5 / (7 * A), 10 / (7 * A)], rtol=1e-6)
    sc1 = oracle.refresh_scale(vol, V, pf, spec, attr, 1.5, None, q, None, None, None, 2, 6, 6,
                               0.0, coord=False)
    np.testing.assert_allclose(sc1, [7.5 / (7 * A)] * 2, rtol=1e-6)


def test_refresh_scale_of_scaled_model_data():
    """Data equal to s_g times the model (CTF . shift(t - offset) . slice, from
    the separately pinned pieces) inside the selected pixels and large noise
    outside them: each group's scale is s_g exactly, so the pixel selection
    (rL <= |k|, rint|k| < rS) leaves out every other pixel; the recentring
    offset is subtracted (data built with t - offset, scored with t and
    offset); the median clamp moves a group 10x off the median to 5 med."""
    N, r, px = 32, 12, 1.5
    rS, rL = 9, 2.5
    P = 12
    rng, vol, V, pf, quat, trans, attr = _sigma_case(P=P, seed=5)
    off = rng.uniform(-1, 1, size=(P, 2))
    model = _model_spectra(vol, V, pf, N, quat, trans - off, attr, r, px)
    i, j, row = oracle.spectrum_pixels(N)
    q2 = i * i + j * j
    v = np.rint(np.hypot(i, j))
    sel = (q2 < rS * rS) & (q2 >= rL * rL) & (v < rS) & (v >= rL)
    grp = np.arange(P) % 3
    s = np.array([0.8, 1.3, 1.1])
    data = 1e3 * (rng.normal(size=model.shape) + 1j * rng.normal(size=model.shape))
    for l in range(P):
        data[l][row[sel], i[sel]] = s[grp[l]] * model[l][row[sel], i[sel]]
    sc = oracle.refresh_scale(vol, V, pf, data, attr, px, np.zeros(P), quat, trans, off, grp, 3,
                              r, rS, rL)
    np.testing.assert_allclose(sc, s, rtol=1e-6)   # FP32 shift phases in the data
    s2 = np.array([0.8, 13.0, 1.1])   # median 1.1: 13 > 5.5 becomes 5.5
    for l in range(P):
        data[l][row[sel], i[sel]] = s2[grp[l]] * model[l][row[sel], i[sel]]
    sc = oracle.refresh_scale(vol, V, pf, data, attr, px, np.zeros(P), quat, trans, off, grp, 3,
                              r, rS, rL)
    np.testing.assert_allclose(sc, [0.8, 5 * np.float32(1.1), 1.1], rtol=1e-6)


def test_norm_correction_hand_worked():
    """Masked data = CTF . shift(t) . slice plus a single residual a_l at one
    pixel inside rL <= |k| < rNorm and large values beyond rNorm: norm_l is
    |a_l|^2 exactly (the residual is X - c M, the pixels outside the range
    are left out, the shift is t without any offset); for |a| = 1, 2, 3, 4
    the median is 6.5 and the factors sqrt(6.5 / norm)."""
    N, r, px = 32, 12, 1.5
    P = 4
    rng, vol, V, pf, quat, trans, attr = _sigma_case(P=P, seed=9)
    model = _model_spectra(vol, V, pf, N, quat, trans, attr, r, px)
    i, j, row = oracle.spectrum_pixels(N)
    rL, rNorm = 2.0, 10.5
    q2 = i * i + j * j
    data = model.copy()
    out = (q2 >= rNorm ** 2) | (q2 < rL ** 2)
    for l in range(P):
        data[l][row[out], i[out]] += 1e4 * (1 + 1j)
        data[l][(N - 5) % N, 4] += (l + 1) * np.exp(1j * 0.3 * l)   # (i, j) = (4, -5)
    norm, m, f = oracle.norm_correction(vol, V, pf, data, attr, px, np.zeros(P), quat, trans, r,
                                        rL, rNorm)
    np.testing.assert_allclose(norm, [1.0, 4.0, 9.0, 16.0], rtol=1e-4)
    assert m == pytest.approx(6.5, rel=1e-4)
    np.testing.assert_allclose(f, np.sqrt(6.5 / np.array([1.0, 4.0, 9.0, 16.0])), rtol=1e-4)
