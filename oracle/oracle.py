"""ctypes binding of the CPU oracle (oracle/thx_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
import this module, and only as the checker.  Each wrapper names the
reference routine its C function restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_thx.so")

_lib = None
_i, _d, _f, _p = C.c_int, C.c_double, C.c_float, C.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    sigs = {
        "thxo_pixel_set": ([_i, _d, _d, _i, _p, _p, _p, _p, _i], _i),
        "thxo_ctf": ([_p, _f, _f, _f, _f, _f, _f, _f, _f, _i, _i, _p, _p, _i], None),
        "thxo_translate": ([_p, _f, _f, _i, _i, _p, _p, _i], None),
        "thxo_rotmat": ([_p, _p], None),
        "thxo_project": ([_p, _p, _i, _p, _p, _p, _i, _i], None),
        "thxo_project_f64": ([_p, _p, _i, _p, _p, _p, _i, _i], None),
        "thxo_logdatavs_pm": ([_p, _p, _p, _p, _i, _i, _p], None),
        "thxo_logdatavs_f64": ([_p, _p, _p, _p, _i], _d),
        "thxo_global_scan_f32": ([_p, _i, _i, _p, _p, _i, _i, _p, _i, _p, _i, _p, _p, _p, _i,
                                  _p, _p, _p, _p, _p, _p, _i], _i),
        "thxo_global_scan_f32_simd256": ([_p, _i, _i, _p, _p, _i, _i, _p, _i, _p, _i, _p, _p,
                                          _p, _i, _p, _p, _p, _p, _p, _p, _i], _i),
        "thxo_logdatavs_pm_simd256": ([_p, _p, _p, _p, _i, _i, _p], _i),
        "thxo_global_scan_f64": ([_p, _i, _i, _p, _p, _i, _i, _p, _i, _p, _i, _p, _p, _p, _i,
                                  _p, _p, _p, _p, _p, _p, _p, _i], _i),
        "thxo_resample": ([_p, _p, _i, _i, _d, _p, _p], _i),
        "thxo_local_scan_f64": ([_p, _i, _i, _p, _p, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p,
                                 _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _i], _i),
        "thxo_blob_ft": ([_p, _i, _i, _p, _p, _i, _d, _p, _p, _p, _i], None),
        "thxo_infer_acg": ([_p, _i, _p, _d, _i, _p], _i),
        "thxo_sym4_top": ([_p, _p, _p], None),
        "thxo_cal_vari_rot": ([_p, _i, _d, _i, _p, _p], _i),
        "thxo_balance_rot": ([_p, _i, _d, _i, _p], _i),
        "thxo_trans_stats": ([_p, _i, _p, _p, _p], None),
        "thxo_keep_peak": ([_p, _i, _d, _i, _p], _d),
        "thxo_post_scan": ([_i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _d, _d, _d, _d,
                            _d, _d, _d, _d, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p], _i),
        "thxo_converge": ([_p, _p, _p, _i, _i, _i], _i),
        "thxo_insert": ([_p, _p, _i, _i, _p, _p, _p, _p, _p, _i, _d], None),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def pixel_set(N, rU, rL=0.0, pf=2):
    """Optimiser::allocPreCalIdx, src/Optimiser.cpp:7942-7992."""
    cap = (N // 2 + 1) * N
    iCol = np.zeros(cap, np.int32)
    iRow = np.zeros(cap, np.int32)
    iSig = np.zeros(cap, np.int32)
    iPxl = np.zeros(cap, np.int32)
    K = load().thxo_pixel_set(N, rU, rL, pf, _ptr(iCol), _ptr(iRow), _ptr(iSig), _ptr(iPxl), cap)
    assert K >= 0
    return iCol[:K].copy(), iRow[:K].copy(), iSig[:K].copy(), iPxl[:K].copy()


def ctf(iCol, iRow, N, pixelSize, voltage, dU, dV, theta, Cs, ampC, phaseShift=0.0):
    """CTF(RFLOAT* dst, ...), src/CTF.cpp:109-145."""
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    out = np.zeros(iCol.size, np.float32)
    load().thxo_ctf(_ptr(out), pixelSize, voltage, dU, dV, theta, Cs, ampC, phaseShift, N, N,
                    _ptr(iCol), _ptr(iRow), iCol.size)
    return out


def translate(tx, ty, N, iCol, iRow):
    """translate(Complex*, ...), src/Image/ImageFunctions.cpp:233-251."""
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    out = np.zeros((iCol.size, 2), np.float32)
    load().thxo_translate(_ptr(out), tx, ty, N, N, _ptr(iCol), _ptr(iRow), iCol.size)
    return out


def rotmat(q):
    """rotate3D(dmat33&, const dvec4&), src/Geometry/Euler.cpp:181-189 (row-major)."""
    q = _c(q, np.float64)
    M = np.zeros(9)
    load().thxo_rotmat(_ptr(M), _ptr(q))
    return M.reshape(3, 3)


def project(vol, V, M, iCol, iRow, pf, f64=False):
    """Projector::project, src/Projector.cpp:356-374 (+ Volume interpolation)."""
    vol = _c(vol, np.float32)
    M = _c(M, np.float64).reshape(9)
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    if f64:
        out = np.zeros((iCol.size, 2), np.float64)
        load().thxo_project_f64(_ptr(out), _ptr(vol), V, _ptr(M), _ptr(iCol), _ptr(iRow),
                                iCol.size, pf)
    else:
        out = np.zeros((iCol.size, 2), np.float32)
        load().thxo_project(_ptr(out), _ptr(vol), V, _ptr(M), _ptr(iCol), _ptr(iRow),
                            iCol.size, pf)
    return out


def logdatavs_pm(dat, pri, ctf_, sig):
    """logDataVSPrior_m_n_huabin, src/Optimiser.cpp:9666-9708; arrays pixel-major [K][n]."""
    dat, pri = _c(dat, np.float32), _c(pri, np.float32)
    ctf_, sig = _c(ctf_, np.float32), _c(sig, np.float32)
    m, n = ctf_.shape
    res = np.zeros(n, np.float32)
    load().thxo_logdatavs_pm(_ptr(dat), _ptr(pri), _ptr(ctf_), _ptr(sig), n, m, _ptr(res))
    return res


def logdatavs_pm_simd256(dat, pri, ctf_, sig):
    """SIMD256Float (src/Optimiser.cpp:8957-9041): pixel-major dat [K][n][2],
    ctf / sig [K][n], pri [K][2] -> [n]."""
    dat, pri = _c(dat, np.float32), _c(pri, np.float32)
    ctf_, sig = _c(ctf_, np.float32), _c(sig, np.float32)
    K, n = ctf_.shape
    out = np.zeros(n, np.float32)
    st = load().thxo_logdatavs_pm_simd256(_ptr(dat), _ptr(pri), _ptr(ctf_), _ptr(sig), n, K,
                                          _ptr(out))
    if st != 0:
        raise RuntimeError("SIMD256Float needs a CPU with AVX2")
    return out


def logdatavs_f64(dat, pri, ctf_, sig):
    dat, pri = _c(dat, np.float32), _c(pri, np.float32)
    ctf_, sig = _c(ctf_, np.float32), _c(sig, np.float32)
    return load().thxo_logdatavs_f64(_ptr(dat), _ptr(pri), _ptr(ctf_), _ptr(sig), ctf_.size)


def global_scan(vol, V, pf, iCol, iRow, N, quat, trans, dat, ctf_, sig, pR, pT,
                f64=True, threads=0, want_dvp=False, simd=False):
    """doGlobalSearch hot loop, src/Optimiser.cpp:711-915 (one class); with
    simd (FP32 only) the likelihood sum is SIMD256Float's (AVX2 required)."""
    vol = _c(vol, np.float32)
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    quat = _c(quat, np.float64).reshape(-1, 4)
    trans = _c(trans, np.float64).reshape(-1, 2)
    dat, ctf_, sig = _c(dat, np.float32), _c(ctf_, np.float32), _c(sig, np.float32)
    pR, pT = _c(pR, np.float64), _c(pT, np.float64)
    K, R, T, P = iCol.size, quat.shape[0], trans.shape[0], ctf_.shape[0]
    if f64:
        wC, wR, wT, base = np.zeros(P), np.zeros((P, R)), np.zeros((P, T)), np.zeros(P)
        dvp = np.zeros((P, R, T)) if want_dvp else None
        st = load().thxo_global_scan_f64(_ptr(vol), V, pf, _ptr(iCol), _ptr(iRow), K, N,
                                         _ptr(quat), R, _ptr(trans), T, _ptr(dat), _ptr(ctf_),
                                         _ptr(sig), P, _ptr(pR), _ptr(pT), _ptr(wC), _ptr(wR),
                                         _ptr(wT), _ptr(base), _ptr(dvp), threads)
        assert st == 0
        return (wC, wR, wT, base, dvp) if want_dvp else (wC, wR, wT, base)
    wC = np.zeros(P, np.float32)
    wR = np.zeros((P, R), np.float32)
    wT = np.zeros((P, T), np.float32)
    base = np.zeros(P, np.float32)
    fn = load().thxo_global_scan_f32_simd256 if simd else load().thxo_global_scan_f32
    st = fn(_ptr(vol), V, pf, _ptr(iCol), _ptr(iRow), K, N, _ptr(quat), R, _ptr(trans), T,
            _ptr(dat), _ptr(ctf_), _ptr(sig), P, _ptr(pR), _ptr(pT), _ptr(wC), _ptr(wR), _ptr(wT),
            _ptr(base), threads)
    if st == -2:
        raise RuntimeError("the SIMD256 scan needs a CPU with AVX2")
    assert st == 0
    return wC, wR, wT, base


def local_scan(vol, V, pf, iCol, iRow, N, quat, nR, trans, nT, pR, pT, dat, ctf_, sig,
               threads=0):
    """One local-search phase (Optimiser::updateSP, src/Optimiser.cpp:1139-1437), FP64.
    quat / pR are the particles' rotations concatenated (counts nR); trans / pT
    likewise (counts nT).  Returns wC[P], wR (concatenated), wT (concatenated),
    base[P], bestR[P], bestT[P]."""
    vol = _c(vol, np.float32)
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    quat = _c(quat, np.float64).reshape(-1, 4)
    trans = _c(trans, np.float64).reshape(-1, 2)
    nR, nT = _c(nR, np.int32), _c(nT, np.int32)
    rOff = np.r_[0, np.cumsum(nR)[:-1]].astype(np.int32)
    tOff = np.r_[0, np.cumsum(nT)[:-1]].astype(np.int32)
    pR, pT = _c(pR, np.float64), _c(pT, np.float64)
    dat, ctf_, sig = _c(dat, np.float32), _c(ctf_, np.float32), _c(sig, np.float32)
    P = len(nR)
    wC, base = np.zeros(P), np.zeros(P)
    wR, wT = np.zeros(int(nR.sum())), np.zeros(int(nT.sum()))
    bR, bT = np.zeros(P, np.int32), np.zeros(P, np.int32)
    st = load().thxo_local_scan_f64(_ptr(vol), V, pf, _ptr(iCol), _ptr(iRow), iCol.size, N,
                                    _ptr(quat), _ptr(rOff), _ptr(nR), _ptr(trans), _ptr(tOff),
                                    _ptr(nT), _ptr(pR), _ptr(pT), _ptr(dat), _ptr(ctf_),
                                    _ptr(sig), P, _ptr(wC), _ptr(wR), _ptr(wT), _ptr(base),
                                    _ptr(bR), _ptr(bT), threads)
    assert st == 0
    return wC, wR, wT, base, bR, bT


def resample(w, u, n_out, u0):
    """Particle::resample(n, pt), src/Particle.cpp:1275-1462, u0 injected, no shuffle."""
    w, u = _c(w, np.float64), _c(u, np.float64)
    idx = np.zeros(n_out, np.int32)
    wn = np.zeros(n_out)
    st = load().thxo_resample(_ptr(w), _ptr(u), w.size, n_out, u0, _ptr(idx), _ptr(wn))
    assert st == 0
    return idx, wn


def blob_ft(V, N, centres, amps, sigma, ix, iy, iz):
    centres = _c(centres, np.float64)
    amps = _c(amps, np.float64)
    ix, iy, iz = _c(ix, np.int32), _c(iy, np.int32), _c(iz, np.int32)
    out = np.zeros((ix.size, 2), np.float32)
    load().thxo_blob_ft(_ptr(out), V, N, _ptr(centres), _ptr(amps), amps.size, sigma,
                        _ptr(ix), _ptr(iy), _ptr(iz), ix.size)
    return out


def infer_acg(q, pre=None, tol=1e-3, max_iter=10000):
    """inferACG(dmat44&, const dmat4&), src/Geometry/DirectionalStat.cpp:87-139.
    Returns (A, iterations)."""
    q = _c(q, np.float64).reshape(-1, 4)
    A = np.zeros(16)
    pre_a = None if pre is None else _c(pre, np.float64)
    it = load().thxo_infer_acg(_ptr(q), q.shape[0], _ptr(pre_a), C.c_double(tol), max_iter,
                               _ptr(A))
    assert it > 0, it
    return A.reshape(4, 4), it


def sym4_top(A):
    v, ev = np.zeros(4), np.zeros(4)
    load().thxo_sym4_top(_ptr(_c(A, np.float64).reshape(16)), _ptr(v), _ptr(ev))
    return v


def cal_vari_rot(q, tol=1e-3, max_iter=10000):
    """Particle::calVari(PAR_R), src/Particle.cpp:1002-1081: (mean, k[3])."""
    q = _c(q, np.float64).reshape(-1, 4)
    mean, k = np.zeros(4), np.zeros(3)
    st = load().thxo_cal_vari_rot(_ptr(q), q.shape[0], C.c_double(tol), max_iter, _ptr(mean),
                                  _ptr(k))
    assert st == 0, st
    return mean, k


def balance_rot(q, tol=1e-3, max_iter=10000):
    """Particle::balanceWeight(PAR_R), src/Particle.cpp:2301-2311 (+ normW)."""
    q = _c(q, np.float64).reshape(-1, 4)
    w = np.zeros(q.shape[0])
    st = load().thxo_balance_rot(_ptr(q), q.shape[0], C.c_double(tol), max_iter, _ptr(w))
    assert st == 0, st
    return w


def trans_stats(t):
    """Means, n-1 standard deviations and balance weights of shifts t[n][2]
    (Particle::calVari(PAR_T) :1083-1097, balanceWeight(PAR_T) :2313-2344)."""
    t = _c(t, np.float64).reshape(-1, 2)
    m, s, w = np.zeros(2), np.zeros(2), np.zeros(t.shape[0])
    load().thxo_trans_stats(_ptr(t), t.shape[0], _ptr(m), _ptr(s), _ptr(w))
    return m, s, w


def keep_peak(u, factor=1e-3, set_factor=False):
    """setPeakFactor(PAR_R) + keepHalfHeightPeak(PAR_R), src/Particle.cpp:1864-1954.
    Returns (u after, factor, index of the first maximum)."""
    u = _c(u, np.float64).copy()
    top = C.c_int(0)
    f = load().thxo_keep_peak(_ptr(u), u.size, C.c_double(factor), int(set_factor),
                              C.byref(top))
    return u, f, top.value


def post_scan(wC, base, wR, wT, pR, pT, grid_q, grid_t, mLR, mLT, u0C, pickC, u0R, u0T,
              min_std_r, min_std_t, pfS, tol=1e-3, max_iter=1000):
    """Post-scan update of one particle, src/Optimiser.cpp:926-1084 (random draws
    given).  wC, base [nCls]; wR [nCls][R]; wT [nCls][T].  Returns a dict with
    cls, idxR, q, wR, idxT, t, wT, k, s, peakR."""
    wC, base = _c(np.atleast_1d(wC), np.float64), _c(np.atleast_1d(base), np.float64)
    nCls = wC.size
    wR, wT = _c(wR, np.float64).reshape(nCls, -1), _c(wT, np.float64).reshape(nCls, -1)
    R, T = wR.shape[1], wT.shape[1]
    pR, pT = _c(pR, np.float64), _c(pT, np.float64)
    gq, gt = _c(grid_q, np.float64).reshape(R, 4), _c(grid_t, np.float64).reshape(T, 2)
    cls = C.c_int(0)
    idxR, idxT = np.zeros(mLR, np.int32), np.zeros(mLT, np.int32)
    q, wRn = np.zeros((mLR, 4)), np.zeros(mLR)
    t, wTn = np.zeros((mLT, 2)), np.zeros(mLT)
    k, sd, pk = np.zeros(3), np.zeros(2), C.c_double(0)
    st = load().thxo_post_scan(nCls, _ptr(wC), _ptr(base), _ptr(wR), _ptr(wT), _ptr(pR),
                               _ptr(pT), _ptr(gq), _ptr(gt), R, T, mLR, mLT, u0C, pickC, u0R,
                               u0T, min_std_r, min_std_t, pfS, tol, max_iter, C.byref(cls),
                               _ptr(idxR), _ptr(q), _ptr(wRn), _ptr(idxT), _ptr(t), _ptr(wTn),
                               _ptr(k), _ptr(sd), C.byref(pk))
    assert st == 0, st
    return dict(cls=cls.value, idxR=idxR, q=q, wR=wRn, idxT=idxT, t=t, wT=wTn, k=k, s=sd,
                peakR=pk.value)


def converge(k, s, sD=None, phase0=1, min_phase=10):
    """Phase at which the local search of one particle stops (-1: not within
    the sequence), src/Optimiser.cpp:1480-1612.  k [n][3]; s [n][2] (s0, s1)
    or [n][3] (s0, s1, rho); sD [n] (defocus deviation; constant by default)."""
    k = _c(k, np.float64).reshape(-1, 3)
    n = k.shape[0]
    s = np.asarray(s, np.float64).reshape(n, -1)
    if s.shape[1] == 2:
        s = np.concatenate([s, np.zeros((n, 1))], axis=1)
    s = _c(s, np.float64)
    sD = _c(np.ones(n) if sD is None else sD, np.float64)
    return load().thxo_converge(_ptr(k), _ptr(s), _ptr(sD), n, phase0, min_phase)


def projectee(vol_ft, N, pf):
    """Projector::setProjectee(Volume, nThread), src/Projector.cpp:123-148, with the
    LINEAR gridCorrection (src/Projector.cpp:524-606), in numpy FP64:
    FFT::bw (c2r, scaled by 1/N^3, src/FFT.cpp:201-230); VOL_PAD_RL (zero padding to
    V = pf N around the origin, negative indices wrapped, include/Image/ImageFunctions.h:
    176-192 with Volume::iRL, include/Image/Volume.h:519-527); division by
    TIK_RL(r / (pf V)) = j0(pi r / (pf V))^2 (src/Functions/Functions.cpp:217-220),
    r = |(i, j, k)| of the padded voxel; FFT::fw (r2c, unnormalised, src/FFT.cpp:172-199).
    vol_ft: complex [N][N][N/2+1] (z, y, x); returns complex [V][V][V/2+1]."""
    V = pf * N
    real = np.fft.irfftn(np.asarray(vol_ft, np.complex128), s=(N, N, N), axes=(0, 1, 2))
    c = np.r_[0:N // 2, -N // 2:0]
    idx = np.where(c >= 0, c, c + V)
    pad = np.zeros((V, V, V))
    pad[np.ix_(idx, idx, idx)] = real
    s = np.r_[0:V // 2, -V // 2:0].astype(np.float64)
    r = np.sqrt(s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2)
    a = np.pi * r / (pf * V)
    j0 = np.ones_like(a)
    nz = a > 0
    j0[nz] = np.sin(a[nz]) / a[nz]
    pad /= j0 ** 2
    return np.fft.rfftn(pad, axes=(0, 1, 2))


EDGE_WIDTH_RL = 6   # include/Macro.h:82


def soft_mask(N, r, ew=EDGE_WIDTH_RL):
    """softMask(Image&, r, ew) (src/Functions/Mask.cpp:333-350) over an N x N
    image with its centre at pixel (N/2, N/2): 1 for u < r,
    0.5 + 0.5 cos(pi (u - r) / ew) for r <= u <= r + ew, 0 beyond;
    u = |(x - N/2, y - N/2)|."""
    h = N // 2
    y, x = np.mgrid[0:N, 0:N]
    u = np.hypot(x - h, y - h)
    m = np.where(u > r + ew, 0.0, 0.5 + 0.5 * np.cos((u - r) / ew * np.pi))
    return np.where(u < r, 1.0, m)


def particle_images(images, bg_radius, iCol, iRow, masked=False):
    """The image path of Optimiser::initImg in numpy FP64, per image: the
    background (centred i^2 + j^2 > r^2) mean and sample standard deviation
    (bgMeanStddev, src/Image/ImageFunctions.cpp:607-620; GSL's sd uses n - 1)
    and (x - mean) / sd (Optimiser::substractBgImg, src/Optimiser.cpp:
    4917-4951); the centre moved to the origin as ImageFile's read stores an
    image (MESH_IMAGE_INDEX, include/Image/ImageFile.h:199-200); the
    unnormalised r2c FFT (Optimiser::fwImg, src/Optimiser.cpp:5003-5013);
    the pixel set gathered at Image::iFTHalf(iCol, iRow)
    (include/Image/Image.h:426-432) as allocPreCal packs datP
    (src/Optimiser.cpp:8005-8016).  With masked, the normalised image is
    multiplied by soft_mask(N, bg_radius) first (maskImg / reMaskImg with
    zeroMask, src/Optimiser.cpp:4953-4985, 6053-6105): the images
    allocPreCal(mask = true) packs.  images: [P][N][N] with the centre at
    pixel (N/2, N/2); returns complex [P][K]."""
    N = np.asarray(images).shape[1]
    iCol = np.asarray(iCol)
    iRow = np.asarray(iRow)
    rows = np.where(iRow >= 0, iRow, iRow + N)
    return image_spectra(images, bg_radius, masked)[:, rows, iCol]


def subtract_bg(images, bg_radius):
    """Optimiser::substractBgImg (src/Optimiser.cpp:4917-4951) in FP64: each
    image minus its background mean over its background sample standard
    deviation (bgMeanStddev, src/Image/ImageFunctions.cpp:607-620), the
    background being the centred pixels with i^2 + j^2 > r^2.  Returns the
    images and the background mask."""
    images = np.asarray(images, np.float64)
    P, N, _ = images.shape
    h = N // 2
    j, i = np.meshgrid(np.arange(N) - h, np.arange(N) - h, indexing="ij")
    bg = (i * i + j * j) > bg_radius ** 2
    out = np.empty_like(images)
    for p in range(P):
        v = images[p][bg]
        out[p] = (images[p] - v.mean()) / v.std(ddof=1)
    return out, bg


def stat_img(images, bg):
    """Optimiser::statImg's noise deviation (src/Optimiser.cpp:4796-4895):
    the mean over the images of bgStddev(0, image, r)
    (src/Image/ImageFunctions.cpp:585-596: GSL's standard deviation of the
    background pixels about the given mean 0, n - 1 in the denominator)."""
    n = int(bg.sum())
    return float(np.mean([np.sqrt(np.sum(x[bg] ** 2) / (n - 1)) for x in images]))


def image_spectra(images, bg_radius, masked=False):
    """The stored half spectra [P][N][N/2 + 1] of the normalised images
    (the steps of particle_images before the gather): substractBgImg,
    statImg, maskImg (masked), normaliseImg (every image times 1 / stdN,
    src/Optimiser.cpp:4987-5001), the centre move and fwImg.  Row r is
    frequency r (r < N/2) or r - N, column c frequency c."""
    y, bg = subtract_bg(images, bg_radius)
    P, N, _ = y.shape
    h = N // 2
    scale = 1.0 / stat_img(y, bg)
    mask = soft_mask(N, bg_radius) if masked else 1.0
    out = np.empty((P, N, h + 1), np.complex128)
    for p in range(P):
        z = np.roll(y[p] * mask * scale, (-h, -h), axis=(0, 1))   # stored[r][c] = file[(r + h) % N][(c + h) % N]
        out[p] = np.fft.rfft2(z)
    return out


def init_sigma(images, bg_radius, n_shell):
    """Optimiser::initSigma (src/Optimiser.cpp:5111-5209) for one group, on
    the normalised unmasked spectra (OPTIMISER_SIGMA_MASK is off,
    include/Config.h:169): the average image avg = mean_p F_p; the average over
    images of powerSpectrum (src/Functions/Spectrum.cpp:161-190: the mean of
    |F|^2 over the pixels of shell u = rint(hypot(i, j)) < r, the loop of
    IMAGE_FOR_EACH_PIXEL_FT, include/Image/Image.h:50-52: column i = 0 .. N/2,
    row j = -N/2 .. N/2 - 1); psAvg(u) = (ringAverage of Re + Im of avg,
    src/Functions/Spectrum.cpp:47-62)^2; sigma^2 = (avgPs - psAvg) / 2 and
    sigRcp = -0.5 / sigma^2.  Returns (sigma^2, sigRcp) for u < n_shell."""
    F = image_spectra(images, bg_radius)
    P, N, W = F.shape
    j = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    i = np.arange(W)[None, :]
    shell = np.rint(np.hypot(i, j)).astype(np.int64)
    avg = F.mean(axis=0)
    pw = (F.real ** 2 + F.imag ** 2).mean(axis=0)
    sig = np.empty(n_shell)
    for u in range(n_shell):
        m = shell == u
        avg_ps = pw[m].mean()
        ps_avg = (avg.real[m] + avg.imag[m]).mean() ** 2
        sig[u] = 0.5 * (avg_ps - ps_avg)
    return sig, -0.5 / sig


def spectrum_pixels(N):
    """The stored half-spectrum pixels in the loop order of
    IMAGE_FOR_EACH_PIXEL_FT (include/Image/Image.h:50-52): column i = 0 .. N/2,
    signed row j = -N/2 .. N/2 - 1.  Returns (i, j, row index into the stored
    spectrum)."""
    j, i = np.meshgrid(np.arange(-N // 2, N // 2), np.arange(N // 2 + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    return i, j, np.where(j >= 0, j, j + N)


def refresh_sigma(vol, V, pf, specM, specN, ctf_attr, pixel_size, cls, quat, trans, offset,
                  group, n_group, r, n_shell, mask_radius):
    """Optimiser::allReduceSigma(mask, group) (src/Optimiser.cpp:6355-6666) in
    FP64, 3-D, OPTIMISER_SIGMA_RANK1ST and OPTIMISER_SIGMA_WHOLE_FREQUENCY on,
    OPTIMISER_SIGMA_MASK off, OPTIMISER_RECENTRE_IMAGE_EACH_ITERATION on
    (include/Config.h:165-198).  Per particle l with best pose (cls, quat,
    trans): the model M = CTF(l) . exp(-2 pi i (k . t) / N) . P_cls(R k) on the
    pixels |k| < r (Projector::project(Image&, rot, t) with maxRadius r,
    src/Projector.cpp:276-296, 456-464; translate(..., r, ...),
    src/Image/ImageFunctions.cpp:158-174; CTF(Image&, ..., r),
    src/CTF.cpp:66-105), zero elsewhere, and M' the same with t - offset.
    With specM / specN the masked / unmasked stored half spectra [P][N][N/2+1]
    and powerSpectrum(u) the mean of |.|^2 over the pixels of shell
    rint(|k|) = u < n_shell (src/Functions/Spectrum.cpp:161-190):
      sigM_g(u) = mean over the group's particles of ps(specM - M)(u) / 2,
      sigN_g(u) = ... ps(specN - M')(u) / 2,
      svd_g(u)  = mean of sqrt(ps(M)(u) / ps(specM)(u)),
      sig = ratio sigM + (1 - ratio) alpha sigN, ratio = min(1, svd),
      alpha = sqrt(pi (maskRadius / (N pixelSize))^2), sigRcp = -0.5 / sig.
    Returns (sig, sigRcp) [n_group][n_shell]."""
    specM, specN = np.asarray(specM), np.asarray(specN)
    P, N = specM.shape[0], specM.shape[1]
    i, j, row = spectrum_pixels(N)
    u = np.rint(np.hypot(i, j)).astype(np.int64)
    keep = u < n_shell
    i, j, row, u = i[keep], j[keep], row[keep], u[keep]
    cnt = np.bincount(u, minlength=n_shell).astype(np.float64)
    inr = i * i + j * j < r * r
    sM = np.zeros((n_group, n_shell))
    sN = np.zeros((n_group, n_shell))
    sv = np.zeros((n_group, n_shell))
    ng = np.zeros(n_group)
    vols = vol if isinstance(vol, (list, tuple)) else [vol]
    for l in range(P):
        a = np.asarray(ctf_attr[l], np.float32)
        c = ctf(i, j, N, pixel_size, a[0], a[1], a[2], a[3], a[4], a[5], a[6]).astype(np.float64)
        Pk = np.zeros(i.size, np.complex128)
        pin = project(vols[int(cls[l])], V, rotmat(quat[l]), i[inr], j[inr], pf, f64=True)
        Pk[inr] = pin[..., 0] + 1j * pin[..., 1]
        t = np.asarray(trans[l], np.float64)
        t2 = t - (np.asarray(offset[l], np.float64) if offset is not None else 0.0)
        M = np.where(inr, c * np.exp(-2j * np.pi * (i * t[0] + j * t[1]) / N) * Pk, 0)
        Mn = np.where(inr, c * np.exp(-2j * np.pi * (i * t2[0] + j * t2[1]) / N) * Pk, 0)
        xm = specM[l][row, i]
        xn = specN[l][row, i]
        ps = lambda z: np.bincount(u, weights=np.abs(z) ** 2, minlength=n_shell) / cnt
        g = int(group[l]) if group is not None else 0
        sM[g] += ps(xm - M) / 2
        sN[g] += ps(xn - Mn) / 2
        sv[g] += np.sqrt(ps(M) / ps(xm))
        ng[g] += 1
    sM /= ng[:, None]
    sN /= ng[:, None]
    sv /= ng[:, None]
    alpha = np.sqrt(np.pi * (mask_radius / (N * pixel_size)) ** 2)
    ratio = np.minimum(1.0, sv)
    sig = ratio * sM + (1 - ratio) * alpha * sN
    with np.errstate(divide="ignore"):
        return sig, -0.5 / sig


def insert(F, T, V, pf, src, ctf_, M, iCol, iRow, w):
    """Reconstructor::insertP (src/Reconstructor.cpp:782-860), trilinear, T
    during the insert, FP64 accumulators F [V][V][V/2+1] complex128 and T
    [V][V][V/2+1] float64, updated in place.  src complex [K], ctf [K], M the
    row-major rotation matrix."""
    assert F.dtype == np.complex128 and T.dtype == np.float64
    assert F.flags.c_contiguous and T.flags.c_contiguous
    s = np.ascontiguousarray(np.stack([np.real(src), np.imag(src)], -1), np.float32)
    c = _c(ctf_, np.float32)
    iCol, iRow = _c(iCol, np.int32), _c(iRow, np.int32)
    load().thxo_insert(_ptr(F), _ptr(T), V, pf, _ptr(s), _ptr(c),
                       _ptr(_c(M, np.float64).reshape(9)), _ptr(iCol), _ptr(iRow), iCol.size, w)


def reconstruct_insert(specN, ctf_attr, pixel_size, cls, quat, trans, offset, w, n_cls, V, pf,
                       rU):
    """The insert loop of Optimiser::reconstructRef (src/Optimiser.cpp:
    7000-7190, 3-D, no CTF search): for each particle and each of its samples
    (quat, trans of shape [P][m][.]), the unmasked packed data at the pixel set
    of radius rU (allocPreCalIdx(rU, 0), allocPreCal(mask = false)) shifted by
    -(t - offset) (translate, src/Image/ImageFunctions.cpp:233-251), its CTF
    from the attributes, inserted with weight w (a number, or one per
    particle) into its class's F and T.
    Returns (F [n_cls][V][V][V/2+1] complex, T, dir [n_cls][3] = sum of
    -R (t - offset, 0) (insertDir), counts [n_cls])."""
    specN = np.asarray(specN)
    P, N = specN.shape[0], specN.shape[1]
    iCol, iRow, _ = pixel_set(N, rU)[:3]
    iCol, iRow = np.asarray(iCol), np.asarray(iRow)
    rows = np.where(iRow >= 0, iRow, iRow + N)
    F = np.zeros((n_cls, V, V, V // 2 + 1), np.complex128)
    T = np.zeros((n_cls, V, V, V // 2 + 1))
    d = np.zeros((n_cls, 3))
    n = np.zeros(n_cls, np.int64)
    quat = np.asarray(quat, np.float64)
    trans = np.asarray(trans, np.float64)
    for l in range(P):
        a = np.asarray(ctf_attr[l], np.float32)
        c = ctf(iCol, iRow, N, pixel_size, *a)
        x = specN[l][rows, iCol]
        for m in range(quat.shape[1]):
            k = int(cls[l][m]) if np.ndim(cls) == 2 else int(cls[l])
            t = trans[l, m] - (np.asarray(offset[l], np.float64) if offset is not None else 0.0)
            s = translate(-t[0], -t[1], N, iCol, iRow)
            src = x * (s[:, 0] + 1j * s[:, 1])
            R = rotmat(quat[l, m])
            insert(F[k], T[k], V, pf, src, c, R, iCol, iRow, float(w[l]) if np.ndim(w) else w)
            d[k] += -(R @ np.array([t[0], t[1], 0.0]))
            n[k] += 1
    return F, T, d, n


def gsl_median(x):
    """median(vec, n) (src/Functions/Functions.cpp:227-233): GSL's quantile
    0.5 of the sorted values, (1 - d) x[lhs] + d x[lhs + 1] with
    lhs = floor((n - 1) / 2), d = (n - 1) / 2 - lhs; returned as RFLOAT
    (float)."""
    s = np.sort(np.asarray(x, np.float64))
    idx = 0.5 * (s.size - 1)
    lhs = int(np.floor(idx))
    d = idx - lhs
    v = s[lhs] if lhs == s.size - 1 else (1 - d) * s[lhs] + d * s[lhs + 1]
    return float(np.float32(v))


def _best_model(vols, V, pf, N, i, j, r, cls_l, q, t, a, pixel_size):
    """CTF and shifted slice of one best pose on the pixels (i, j): the slice
    P_cls(R k) on |k| < r (zero beyond), times exp(-2 pi i k.t / N) when t is
    given (Projector::project(Image&, rot, t), src/Projector.cpp:456-464), and
    the CTF from the attributes a (CTF(Image&, ...), src/CTF.cpp:66-105)."""
    inr = i * i + j * j < r * r
    Pk = np.zeros(i.size, np.complex128)
    pin = project(vols[cls_l], V, rotmat(q), i[inr], j[inr], pf, f64=True)
    Pk[inr] = pin[..., 0] + 1j * pin[..., 1]
    if t is not None:
        Pk *= np.exp(-2j * np.pi * (i * t[0] + j * t[1]) / N)
    c = ctf(i, j, N, pixel_size, a[0], a[1], a[2], a[3], a[4], a[5], a[6]).astype(np.float64)
    return Pk, c


def refresh_scale(vol, V, pf, specN, ctf_attr, pixel_size, cls, quat, trans, offset, group,
                  n_group, r, rS, rL, coord=True):
    """Optimiser::refreshScale(coord, group) (src/Optimiser.cpp:5715-6023) in
    FP64, 3-D, with the switches of include/Config.h: OPTIMISER_SCALE_MASK
    off (the unmasked images), OPTIMISER_RECENTRE_IMAGE_EACH_ITERATION on
    (translation t - offset), OPTIMISER_REFRESH_SCALE_SPECTRUM off (one ratio
    of sums per group), OPTIMISER_REFRESH_SCALE_ZERO_FREQ_NO_COORD on (without
    coordinates: class 0, no translation, rS = 1, rL = 0).  rS is clamped to r.
    scaleDataVSPrior (:10021-10063) over the loop IMAGE_FOR_PIXEL_R_FT(rS + 1)
    (every stored pixel, both halves of the i = 0 column) with
    rL^2 <= i^2 + j^2 < rS^2 and rL <= rint|k| < rS adds, with the plain complex
    products of the code (not X conj(M)),
        sXA += Re(X . M) c,   sAA += Re(M . M) c^2
    for the data X, the shifted slice M and the CTF c.  Per group
    (group None: every particle in one sum, the same scale for all groups)
    scale = sum sXA / sum sAA; then with med = median(scale) (float), a scale
    above |5 med| in magnitude becomes 5 med and one below |med / 5| becomes
    med / 5.  Returns the n_group scales."""
    specN = np.asarray(specN)
    P, N = specN.shape[0], specN.shape[1]
    vols = vol if isinstance(vol, (list, tuple)) else [vol]
    if not coord:
        rS, rL = 1, 0.0
    rS = min(rS, r)
    i, j, row = spectrum_pixels(N)
    q2 = (i * i + j * j).astype(np.float64)
    v = np.rint(np.hypot(i, j))
    keep = (q2 < float(rS) ** 2) & (q2 >= float(rL) ** 2) & (v < rS) & (v >= rL)
    i, j, row = i[keep], j[keep], row[keep]
    xa = np.zeros(n_group)
    aa = np.zeros(n_group)
    for l in range(P):
        a = np.asarray(ctf_attr[l], np.float32)
        if coord:
            t = np.asarray(trans[l], np.float64) - (np.asarray(offset[l], np.float64)
                                                    if offset is not None else 0.0)
            M, c = _best_model(vols, V, pf, N, i, j, r, int(cls[l]) if cls is not None else 0,
                               quat[l], t, a, pixel_size)
        else:
            M, c = _best_model(vols, V, pf, N, i, j, r, 0, quat[l], None, a, pixel_size)
        x = specN[l][row, i]
        g = int(group[l]) if group is not None else 0
        xa[g] += np.sum(np.real(x * M) * c)
        aa[g] += np.sum(np.real(M * M) * c * c)
    if group is None:
        scale = np.full(n_group, xa[0] / aa[0])
    else:
        scale = xa / aa
    med = gsl_median(scale)
    out = scale.copy()
    for g in range(n_group):
        if abs(scale[g]) > abs(med * 5):
            out[g] = med * 5
        elif abs(scale[g]) < abs(med / 5):
            out[g] = med / 5
    return out


def norm_correction(vol, V, pf, specM, ctf_attr, pixel_size, cls, quat, trans, r, rL, rNorm):
    """Optimiser::normCorrection (src/Optimiser.cpp:6161-6351) in FP64, 3-D,
    no CTF search, OPTIMISER_NORM_MASK on (the masked images, translation t
    without the offset): per particle the power of X_masked - c M over the
    stored pixels (IMAGE_FOR_EACH_PIXEL_FT) with rL^2 <= i^2 + j^2 < rNorm^2,
    M the slice P_cls(R k) on |k| < r shifted by t, c the CTF; m = median of
    the powers (float); every image is then multiplied by sqrt(m / norm).
    Returns (norm [P], m, factor [P])."""
    specM = np.asarray(specM)
    P, N = specM.shape[0], specM.shape[1]
    vols = vol if isinstance(vol, (list, tuple)) else [vol]
    rNorm = min(float(rNorm), float(r))
    i, j, row = spectrum_pixels(N)
    q2 = (i * i + j * j).astype(np.float64)
    keep = (q2 >= float(rL) ** 2) & (q2 < rNorm ** 2)
    i, j, row = i[keep], j[keep], row[keep]
    norm = np.zeros(P)
    for l in range(P):
        a = np.asarray(ctf_attr[l], np.float32)
        M, c = _best_model(vols, V, pf, N, i, j, r, int(cls[l]) if cls is not None else 0,
                           quat[l], np.asarray(trans[l], np.float64), a, pixel_size)
        x = specM[l][row, i]
        norm[l] = np.sum(np.abs(x - c * M) ** 2)
    m = gsl_median(norm)
    return norm, m, np.sqrt(m / norm)
