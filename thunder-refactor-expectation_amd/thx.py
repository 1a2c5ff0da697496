"""ctypes binding of libthunder_expect.so (the C ABI in include/thunder_expect.h).

Python is only the test and benchmark driver: every computation goes through
the HIP library.  Loading fails loudly when the library is missing; there is
no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libthunder_expect.so")

FP32, BF16, FP16, FP8 = 0, 1, 2, 3
PRECISIONS = {"fp32": FP32, "bf16": BF16, "fp16": FP16, "fp8": FP8}

_lib = None

_i = C.c_int
_d = C.c_double
_p = C.c_void_p
_u64 = C.c_uint64

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "thx_last_error": [],
    "thx_device_count": [_p],
    "thx_ExpectRotran": [_p, _p, _p, _p, _p, _p, _i, _i, _i, _i],
    "thx_ExpectProject": [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i],
    "thx_ExpectGlobal3D": [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                           _i, _i, _i, _i, _i, _i],
    "thx_ctx_create": [_i, _p],
    "thx_ctx_destroy": [_p],
    "thx_ctx_set_pixels": [_p, _p, _p, _p, _i, _i, _i],
    "thx_ctx_set_volume": [_p, _i, _p, _i],
    "thx_ctx_get_volume": [_p, _i, _p],
    "thx_ctx_project": [_p, _i, _i, _p, _i, _p, _p, _i, _p],
    "thx_ctx_set_projectee": [_p, _i, _p, _i, _i, _p],
    "thx_ctx_set_particle_images": [_p, _p, _i, _i, C.c_float, _i, _p, _p, _p],
    "thx_ctx_init_sigma": [_p, _i, _i, _p, _p, _p],
    "thx_ctx_get_image_stats": [_p, _p, _p],
    "thx_ctx_insert": [_p, _p, _p, _p, _p, _i, _p, _i, _i, _i, C.c_float, _p],
    "thx_ctx_get_reconstruction": [_p, _i, _p, _p, _p, _p],
    "thx_ctx_refresh_sigma": [_p, _p, _p, _p, _p, _p, _i, _i, _i, C.c_float, C.c_float, _i, _p,
                              _p, _p],
    "thx_ctx_refresh_scale": [_p, _i, _p, _p, _p, _p, _p, _i, _i, _i, C.c_float, C.c_float, _i,
                              _p, _p],
    "thx_ctx_norm_correction": [_p, _p, _p, _p, _i, C.c_float, C.c_float, C.c_float, _i, _p, _p,
                                _p],
    "thx_ctx_set_particles": [_p, _p, _p, _p, _i],
    "thx_ctx_set_ctf": [_p, _p, _i, C.c_float],
    "thx_ctx_get_particles": [_p, _p, _p, _p],
    "thx_ctx_set_grid": [_p, _p, _i, _p, _i, _p, _p],
    "thx_ctx_score_global": [_p, _i, _i],
    "thx_ctx_get_global": [_p, _i, _p, _p, _p, _p],
    "thx_ctx_get_best": [_p, _i, _p, _p, _p],
    "thx_ctx_score_one": [_p, _i, _i, _p],
    "thx_ctx_sync": [_p],
    "thx_ctx_set_local_options": [_p, _i, _i],
    "thx_ctx_local_mode": [_p, _p],
    "thx_ctx_set_local_gather": [_p, _i],
    "thx_ctx_set_select_order": [_p, _i],
    "thx_ctx_set_learned_fallback_selection": [_p, _i, _i],
    "thx_ctx_score_local": [_p, _p, _p, _p, _p, _p, _p, _p, _i],
    "thx_ctx_get_local": [_p, _p, _p, _p, _p, _p, _p],
    "thx_ctx_last_timing": [_p, _p, _i, _p, _i, _p],
    "thx_pf_resample": [_p, _p, _p, _p, _p, _i, _u64, _p, _p, _p, _p],
    "thx_pf_resample_noalloc": [_p, _p, _p, _p, _p, _i, _u64, _p, _p, _p, _p],
    "thx_pf_rot_mean": [_p, _p, _p, _p, _i, _p, _p],
    "thx_pf_perturb": [_p, _p, _p, _p, _p, _i, _p, _p, _p, _d, _u64, _i, _p],
    "thx_proposal_sample": [_p, _p, _i, _i, _i, _i, _u64, _p, _p, _p],
    "thx_cnn_create": [_i, _i, _u64, _p],
    "thx_cnn_destroy": [_p],
    "thx_cnn_num_params": [_p, _p],
    "thx_cnn_get_params": [_p, _p],
    "thx_cnn_set_params": [_p, _p],
    "thx_cnn_forward": [_p, _p, _i, _p],
    "thx_cnn_loss_grad": [_p, _p, _i, _p, _p, _p, _p],
    "thx_cnn_train_step": [_p, _p, _i, _p, _p, C.c_float, _p],
    "thx_ctx_images": [_p, _p],
    "thx_ctx_images_prec": [_p, _i, _p],
    "thx_cnn_train_ctx": [_p, _p, _p, _p, _i, _i, C.c_float, _p],
    "thx_cnn_propose_ctx": [_p, _p, _p, _p],
    "thx_cnn_components": [_p, _p],
    "thx_cnn_set_precision": [_p, C.c_int],
    "thx_ctx_expect_learned": [_p, _p, _i, _i, _i, _d, _d, _d, _d, _i, _p],
    "thx_ctx_expect_seeded": [_p, _p, _i, _i, _i, _i, _d, _i, _d, _d, _d, _d, _p, _i, _u64, _p,
                              _p],
    "thx_ctx_get_learned": [_p, _p, _p, _p, _p, _p, _p, _p],
    "thx_pf_set_state": [_p, _i, _i, _p, _p, _p, _p, _p, _p],
    "thx_pf_from_global": [_p, _i, _i, _i, _d, _d, _d, _p, _u64],
    "thx_pf_from_global_classes": [_p, _i, _i, _i, _d, _d, _d, _p, _u64, _p],
    "thx_pf_from_global_given": [_p, _i, _i, _i, _d, _d, _d, _p, _p, _p, _p, _p, _p],
    "thx_pf_converge": [_p, _p, _p, _i, _i, _i, _p],
    "thx_pf_mat4": [_p, _p, _i, _p, _p, _p],
    "thx_pf_refine": [_p, _p, _i, _i, _d, _d, _i, _u64, _p],
    "thx_pf_get_state": [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "thx_pf_rot_stats": [_p, _p, _i, _i, _p, _p, _p, _p, _p],
    "thx_pf_trans_stats": [_p, _p, _i, _i, _p, _p, _p],
    "thx_pf_keep_peak": [_p, _p, _i, _i, _p, _i, _p],
    "thx_ctx_synth_volume": [_p, _i, _i, _i, _i, _d, _d, _u64, _p],
    "thx_ctx_synth_volume_blobs": [_p, _i, _i, _i, _i, _d, _p, _p],
    "thx_ctx_synth_particles": [_p, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d,
                                _u64, _p, _p, _p],
}
_SIGS.update({
    "thx_thu_read": [C.c_char_p, _p, _i, _p],
    "thx_thu_write": [C.c_char_p, _p, _i, _i],
    "thx_mrc_info": [C.c_char_p, _p, _p, _p, _p],
    "thx_mrc_read": [C.c_char_p, _i, _i, _p],
    "thx_mrc_write": [C.c_char_p, _p, _i, _i, _i, C.c_float],
    "thx_ctx_load_thu": [_p, C.c_char_p, C.c_char_p, _i, _i, C.c_float, C.c_float, _i, _p],
})
_RESTYPES = {"thx_last_error": C.c_char_p}


class ThuRecord(C.Structure):
    """One .thu line (thx_thu_record in thunder_expect.h)."""
    _fields_ = [("ctf", C.c_float * 7), ("particle", C.c_char * 256),
                ("micrograph", C.c_char * 256), ("coord", C.c_double * 2), ("group", C.c_int),
                ("cls", C.c_int), ("quat", C.c_double * 4), ("k", C.c_double * 3),
                ("trans", C.c_double * 2), ("stdTrans", C.c_double * 2), ("df", C.c_double),
                ("stdDf", C.c_double), ("score", C.c_double)]


def thu_read(path):
    """The particles of a .thu file as a list of ThuRecord."""
    n = C.c_int(0)
    check(load().thx_thu_read(path.encode(), None, 0, C.byref(n)), "thu_read")
    arr = (ThuRecord * n.value)()
    check(load().thx_thu_read(path.encode(), C.cast(arr, C.c_void_p), n.value, C.byref(n)),
          "thu_read")
    return list(arr)


def thu_write(path, records, append=False):
    arr = (ThuRecord * len(records))(*records)
    check(load().thx_thu_write(path.encode(), C.cast(arr, C.c_void_p), len(records),
                               int(bool(append))), "thu_write")


def mrc_info(path):
    v = [C.c_int(0) for _ in range(4)]
    check(load().thx_mrc_info(path.encode(), *[C.byref(x) for x in v]), "mrc_info")
    return tuple(x.value for x in v)   # nx, ny, nz, mode


def mrc_read(path, first=0, count=None):
    nx, ny, nz, _ = mrc_info(path)
    count = nz - first if count is None else count
    out = np.zeros((count, ny, nx), np.float32)
    check(load().thx_mrc_read(path.encode(), first, count, ptr(out)), "mrc_read")
    return out


def mrc_write(path, data, pixel_size=1.0):
    data = c32(data)
    if data.ndim == 2:
        data = data[None]
    nz, ny, nx = data.shape
    check(load().thx_mrc_write(path.encode(), ptr(data), nx, ny, nz, float(pixel_size)),
          "mrc_write")


class THXError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise THXError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def check(status, what=""):
    if status != 0:
        msg = load().thx_last_error().decode(errors="replace")
        raise THXError(f"{what} failed with status {status}: {msg}")


def c32(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def particle_arrays(P, K, ctf, sig, dat=None):
    """The per-particle arrays of set_particles / set_particle_images as
    contiguous float32, checked against (P, K) (dat: (P, K) complex as
    (P, K, 2) floats).  The library copies P * K values from each host
    pointer, so a shorter array would be read past its end: raise instead."""
    ctf = None if ctf is None else c32(ctf)
    sig = None if sig is None else c32(sig)
    for name, a in (("ctf", ctf), ("sig", sig)):
        if a is not None and a.shape != (P, K):
            raise ValueError(f"{name} has shape {a.shape}, expected {(P, K)}")
    if dat is not None:
        dat = c32(dat)
        if dat.size != 2 * P * K or dat.shape[:2] != (P, K):
            raise ValueError(f"dat has shape {dat.shape}, expected {(P, K, 2)}")
    return ctf, sig, dat


def device_count():
    n = C.c_int(0)
    st = load().thx_device_count(C.byref(n))
    return n.value if st == 0 else 0


# ----------------------------------------------------------------------------
# Reference-shaped entry points
# ----------------------------------------------------------------------------

def expect_rotran(trans, rot, iCol, iRow, idim):
    trans = c32(trans, np.float64).reshape(-1, 2)
    rot = c32(rot, np.float64).reshape(-1, 4)
    iCol, iRow = c32(iCol, np.int32), c32(iRow, np.int32)
    nR, nT, K = rot.shape[0], trans.shape[0], iCol.size
    traP = np.zeros((nT, K, 2), np.float32)
    rotMat = np.zeros((nR, 9), np.float64)
    check(load().thx_ExpectRotran(ptr(traP), ptr(trans), ptr(rot), ptr(rotMat), ptr(iCol),
                                  ptr(iRow), nR, nT, idim, K), "ExpectRotran")
    return traP, rotMat


def expect_project(volume, rotMat, iCol, iRow, pf, vdim, interp=1):
    volume = c32(volume)
    rotMat = c32(rotMat, np.float64).reshape(-1, 9)
    iCol, iRow = c32(iCol, np.int32), c32(iRow, np.int32)
    nR, K = rotMat.shape[0], iCol.size
    rotP = np.zeros((nR, K, 2), np.float32)
    check(load().thx_ExpectProject(ptr(volume), ptr(rotP), ptr(rotMat), ptr(iCol), ptr(iRow),
                                   nR, pf, interp, vdim, K), "ExpectProject")
    return rotP


def expect_global3d(rotP, traP, datP, ctfP, sigRcpP, pR, pT, kIdx, nK, wC, wR, wT, baseL):
    """In-place update of wC/wR/wT/baseL (float32 arrays), as ExpectGlobal3D."""
    rotP, traP, datP = c32(rotP), c32(traP), c32(datP)
    ctfP, sigRcpP = c32(ctfP), c32(sigRcpP)
    pR, pT = c32(pR, np.float64), c32(pT, np.float64)
    nR, nT = pR.size, pT.size
    P = ctfP.shape[0]
    K = ctfP.shape[1]
    for a in (wC, wR, wT, baseL):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    check(load().thx_ExpectGlobal3D(ptr(rotP), ptr(traP), ptr(datP), ptr(ctfP), ptr(sigRcpP),
                                    ptr(wC), ptr(wR), ptr(wT), ptr(pR), ptr(pT), ptr(baseL),
                                    kIdx, nK, nR, nT, K, P), "ExpectGlobal3D")


# ----------------------------------------------------------------------------
# Context API
# ----------------------------------------------------------------------------

class Context:
    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(load().thx_ctx_create(device, C.byref(self._h)), "ctx_create")
        self.K = 0
        self.P = 0
        self.R = 0
        self.T = 0

    def close(self):
        if self._h:
            load().thx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_pixels(self, iCol, iRow, iSig, idim, pf):
        iCol, iRow = c32(iCol, np.int32), c32(iRow, np.int32)
        iSig = None if iSig is None else c32(iSig, np.int32)
        check(load().thx_ctx_set_pixels(self._h, ptr(iCol), ptr(iRow), ptr(iSig), iCol.size,
                                        idim, pf), "set_pixels")
        self.K = iCol.size
        self.idim, self.pf = idim, pf

    def set_volume(self, cls, vol, vdim):
        vol = c32(vol)
        check(load().thx_ctx_set_volume(self._h, cls, ptr(vol), vdim), "set_volume")

    def set_projectee(self, cls, vol_ft, N, pf=2):
        """Projectee preparation on the GPU (Projector::setProjectee): vol_ft is
        the N^3 map's half-complex FT [N][N][N/2+1]; class cls then holds the
        padded, grid-corrected (pf N)^3 projectee.  Returns the inverse FFT,
        pad + correction and forward FFT times (ms)."""
        ft = np.ascontiguousarray(np.asarray(vol_ft).astype(np.complex64))
        assert ft.shape == (N, N, N // 2 + 1), ft.shape
        ms = np.zeros(3, np.float32)
        check(load().thx_ctx_set_projectee(self._h, cls, ptr(ft), N, pf, ptr(ms)),
              "set_projectee")
        return ms

    def set_particle_images(self, images, bg_radius, ctf, sig, soft_mask=False):
        """Particle images [P][N][N] (centre at pixel (N/2, N/2)) normalised by
        their background outside bg_radius, FFT'd and packed into the pixel set
        on the GPU (Optimiser::initImg's image path + allocPreCal's datP);
        with soft_mask, the packed data are those of the images times the
        soft mask of radius bg_radius (6-pixel cosine edge).  ctf and sig
        [P][K] as for set_particles.  Returns the normalisation, FFT and
        packing times (ms)."""
        img = c32(images)
        if img.ndim != 3 or img.shape[1] != img.shape[2]:
            raise ValueError(f"images have shape {img.shape}, expected (P, N, N)")
        P, N = img.shape[0], img.shape[1]
        ctf, sig, _ = particle_arrays(P, self.K, ctf, sig)
        ms = np.zeros(3, np.float32)
        check(load().thx_ctx_set_particle_images(self._h, ptr(img), P, N, float(bg_radius),
                                                  int(bool(soft_mask)), ptr(ctf),
                                                  ptr(sig), ptr(ms)),
              "set_particle_images")
        self.P = P
        return ms

    def refresh_sigma(self, quat, trans, r, n_shell, mask_radius, pixel_size, cls=None,
                      offset=None, group=None, n_group=1, apply=False):
        """sigma^2 and sigRcp [n_group][n_shell] refreshed from the particles'
        best poses (Optimiser::allReduceSigma); returns (sig, sigRcp, ms)."""
        P = self.P
        quat = c32(quat, np.float64).reshape(P, 4)
        trans = c32(trans, np.float64).reshape(P, 2)
        arr = lambda a, dt, n: None if a is None else c32(a, dt).reshape(P, n) if n > 1 else c32(a, dt)
        sig = np.zeros((n_group, n_shell))
        rcp = np.zeros((n_group, n_shell))
        ms = np.zeros(1, np.float32)
        check(load().thx_ctx_refresh_sigma(self._h, ptr(arr(cls, np.int32, 1)), ptr(quat),
                                            ptr(trans), ptr(arr(offset, np.float64, 2)),
                                            ptr(arr(group, np.int32, 1)), n_group, r, n_shell,
                                            float(mask_radius), float(pixel_size), int(bool(apply)),
                                            ptr(sig), ptr(rcp), ptr(ms)), "refresh_sigma")
        return sig, rcp, float(ms[0])

    def refresh_scale(self, quat, trans, r, rS, rL, pixel_size, cls=None, offset=None,
                      group=None, n_group=1, coord=True, apply=0):
        """Intensity scale per group from the best poses (Optimiser::refreshScale,
        correctScale with apply 1 (images) or 2 (volumes)); returns (scale, ms)."""
        P = self.P
        q = None if quat is None else c32(quat, np.float64).reshape(P, 4)
        t = None if trans is None else c32(trans, np.float64).reshape(P, 2)
        o = None if offset is None else c32(offset, np.float64).reshape(P, 2)
        cl = None if cls is None else c32(cls, np.int32)
        g = None if group is None else c32(group, np.int32)
        scale = np.zeros(n_group)
        ms = np.zeros(1, np.float32)
        check(load().thx_ctx_refresh_scale(self._h, int(bool(coord)), ptr(cl), ptr(q), ptr(t),
                                            ptr(o), ptr(g), int(n_group), int(r), int(rS),
                                            float(rL), float(pixel_size), int(apply), ptr(scale),
                                            ptr(ms)), "refresh_scale")
        return scale, float(ms[0])

    def norm_correction(self, quat, trans, r, rL, r_norm, pixel_size, cls=None, apply=False):
        """Noise normalisation from the best poses (Optimiser::normCorrection);
        returns (norm [P], median, ms)."""
        P = self.P
        q = c32(quat, np.float64).reshape(P, 4)
        t = c32(trans, np.float64).reshape(P, 2)
        cl = None if cls is None else c32(cls, np.int32)
        norm = np.zeros(P)
        med = np.zeros(1)
        ms = np.zeros(1, np.float32)
        check(load().thx_ctx_norm_correction(self._h, ptr(cl), ptr(q), ptr(t), int(r), float(rL),
                                              float(r_norm), float(pixel_size), int(bool(apply)),
                                              ptr(norm), ptr(med), ptr(ms)), "norm_correction")
        return norm, float(med[0]), float(ms[0])

    def insert(self, quat, trans, r_u, pixel_size, cls=None, offset=None, w=None, n_cls=1,
               clear=True):
        """Reconstruction insert of every particle's m samples (quat [P][m][4],
        trans [P][m][2], cls [P][m]); returns the kernel time in ms."""
        P = self.P
        quat = c32(quat, np.float64).reshape(P, -1, 4)
        m = quat.shape[1]
        trans = c32(trans, np.float64).reshape(P, m, 2)
        cl = None if cls is None else c32(np.broadcast_to(np.asarray(cls).reshape(P, -1), (P, m)),
                                          np.int32)
        off = None if offset is None else c32(offset, np.float64).reshape(P, 2)
        ww = None if w is None else c32(w, np.float64)
        ms = np.zeros(1, np.float32)
        check(load().thx_ctx_insert(self._h, ptr(cl), ptr(quat), ptr(trans), ptr(off), m, ptr(ww),
                                    int(r_u), int(n_cls), int(bool(clear)), float(pixel_size),
                                    ptr(ms)), "insert")
        return float(ms[0])

    def get_reconstruction(self, cls=0):
        """(F complex [V][V][V/2+1], T [V][V][V/2+1], dir[3], count)."""
        V = self.idim * self.pf
        F = np.zeros((V, V, V // 2 + 1), np.complex128)
        T = np.zeros((V, V, V // 2 + 1))
        d = np.zeros(3)
        n = np.zeros(1)
        check(load().thx_ctx_get_reconstruction(self._h, cls, ptr(F), ptr(T), ptr(d), ptr(n)),
              "get_reconstruction")
        return F, T, d, float(n[0])

    def image_stats(self):
        """(stdN, scale) of the last set_particle_images call: statImg's mean
        background deviation and normaliseImg's scale 1 / stdN."""
        sN = np.zeros(1, np.float64)
        sc = np.zeros(1, np.float32)
        check(load().thx_ctx_get_image_stats(self._h, ptr(sN), ptr(sc)), "image_stats")
        return float(sN[0]), float(sc[0])

    def init_sigma(self, n_shell, apply=False):
        """Initial sigma^2 and sigRcp = -0.5 / sigma^2 per shell (0 .. n_shell - 1)
        from the last set_particle_images call (Optimiser::initSigma); with
        apply, the particles' sigRcp becomes sigRcp[iSig].  Returns
        (sigma^2, sigRcp, kernel ms)."""
        sig = np.zeros(n_shell, np.float64)
        rcp = np.zeros(n_shell, np.float64)
        ms = np.zeros(1, np.float32)
        check(load().thx_ctx_init_sigma(self._h, int(n_shell), int(bool(apply)), ptr(sig),
                                         ptr(rcp), ptr(ms)), "init_sigma")
        return sig, rcp, float(ms[0])

    def project(self, cls, rot_mat, iCol, iRow, pf=2):
        """Slices of the resident class volume for column-major 3x3 matrices
        rot_mat [nR][9] at the pixels (iCol, iRow): [nR][nPxl][2] float."""
        m = c32(rot_mat, np.float64).reshape(-1, 9)
        iCol, iRow = c32(iCol, np.int32), c32(iRow, np.int32)
        out = np.zeros((m.shape[0], iCol.size, 2), np.float32)
        check(load().thx_ctx_project(self._h, cls, pf, ptr(m), m.shape[0], ptr(iCol), ptr(iRow),
                                     iCol.size, ptr(out)), "project")
        return out

    def get_volume(self, cls, vdim):
        out = np.zeros((vdim, vdim, vdim // 2 + 1, 2), np.float32)
        check(load().thx_ctx_get_volume(self._h, cls, ptr(out)), "get_volume")
        return out

    def load_thu(self, thu_path, par_prefix, first, count, pixel_size, bg_radius,
                 soft_mask=False):
        """Particles first .. first + count - 1 of a .thu file (images from
        their MRC stacks under par_prefix, CTF from their attributes)."""
        ms = np.zeros(3, np.float32)
        check(load().thx_ctx_load_thu(self._h, thu_path.encode(), par_prefix.encode(), first,
                                      count, float(pixel_size), float(bg_radius),
                                      int(bool(soft_mask)), ptr(ms)), "load_thu")
        self.P = count
        return ms

    def set_ctf(self, attr, pixel_size):
        """The particles' CTF from their attributes attr [P][7] = (voltage,
        defocusU, defocusV, defocusTheta, Cs, amplitudeContrast, phaseShift),
        evaluated on the device (CTF(RFLOAT*), src/CTF.cpp:109-145)."""
        attr = c32(attr).reshape(-1, 7)
        check(load().thx_ctx_set_ctf(self._h, ptr(attr), attr.shape[0], float(pixel_size)),
              "set_ctf")

    def set_particles(self, dat, ctf, sig):
        P = np.shape(dat)[0]
        ctf, sig, dat = particle_arrays(P, self.K, ctf, sig, dat)
        check(load().thx_ctx_set_particles(self._h, ptr(dat), ptr(ctf), ptr(sig), P),
              "set_particles")
        self.P = P

    def get_particles(self):
        dat = np.zeros((self.P, self.K, 2), np.float32)
        ctf = np.zeros((self.P, self.K), np.float32)
        sig = np.zeros((self.P, self.K), np.float32)
        check(load().thx_ctx_get_particles(self._h, ptr(dat), ptr(ctf), ptr(sig)),
              "get_particles")
        return dat, ctf, sig

    def set_grid(self, quat, trans, pR=None, pT=None):
        quat = c32(quat, np.float64).reshape(-1, 4)
        trans = c32(trans, np.float64).reshape(-1, 2)
        R, T = quat.shape[0], trans.shape[0]
        pR = np.full(R, 1.0 / R) if pR is None else c32(pR, np.float64)
        pT = np.full(T, 1.0 / T) if pT is None else c32(pT, np.float64)
        check(load().thx_ctx_set_grid(self._h, ptr(quat), R, ptr(trans), T, ptr(pR), ptr(pT)),
              "set_grid")
        self.R, self.T = R, T
        self.pR, self.pT = pR, pT
        self.grid_quat, self.grid_trans = quat, trans

    def score_global(self, cls=0, precision=FP32):
        check(load().thx_ctx_score_global(self._h, cls, precision), "score_global")

    def sync(self):
        check(load().thx_ctx_sync(self._h), "sync")

    def get_global(self, nK=1):
        P, R, T = self.P, self.R, self.T
        wC = np.zeros((P, nK))
        wR = np.zeros((P, nK, R))
        wT = np.zeros((P, nK, T))
        base = np.zeros(P)
        check(load().thx_ctx_get_global(self._h, nK, ptr(wC), ptr(wR), ptr(wT), ptr(base)),
              "get_global")
        return wC, wR, wT, base

    def get_best(self, cls=0):
        bR = np.zeros(self.P, np.int32)
        bT = np.zeros(self.P, np.int32)
        sc = np.zeros(self.P)
        check(load().thx_ctx_get_best(self._h, cls, ptr(bR), ptr(bT), ptr(sc)), "get_best")
        return bR, bT, sc

    def score_one(self, cls, p):
        out = np.zeros((self.R, self.T), np.float32)
        check(load().thx_ctx_score_one(self._h, cls, p, ptr(out)), "score_one")
        return out

    def score_local(self, quat, nR, trans, nT, pR=None, pT=None, cls=None, precision=FP32):
        quat = c32(quat, np.float64).reshape(-1, 4)
        trans = c32(trans, np.float64).reshape(-1, 2)
        nR, nT = c32(nR, np.int32), c32(nT, np.int32)
        if pR is None:
            pR = np.concatenate([np.full(n, 1.0 / n) for n in nR])
        if pT is None:
            pT = np.concatenate([np.full(n, 1.0 / n) for n in nT])
        pR, pT = c32(pR, np.float64), c32(pT, np.float64)
        cls = None if cls is None else c32(cls, np.int32)
        check(load().thx_ctx_score_local(self._h, ptr(cls), ptr(quat), ptr(nR), ptr(trans),
                                         ptr(nT), ptr(pR), ptr(pT), precision), "score_local")
        self._loc = (int(nR.sum()), int(nT.sum()))

    def get_local(self):
        Rt, Tt = self._loc
        wC, base = np.zeros(self.P), np.zeros(self.P)
        wR, wT = np.zeros(Rt), np.zeros(Tt)
        bR, bT = np.zeros(self.P, np.int32), np.zeros(self.P, np.int32)
        check(load().thx_ctx_get_local(self._h, ptr(wC), ptr(wR), ptr(wT), ptr(base), ptr(bR),
                                       ptr(bT)), "get_local")
        return wC, wR, wT, base, bR, bT

    def set_local_options(self, mode=0, graph=False):
        """Local scorer launch shape (0 autotuned, 1 direct, 2 persistent) and
        hipGraph capture of the local phase."""
        check(load().thx_ctx_set_local_options(self._h, mode, int(graph)), "set_local_options")

    def local_mode(self):
        """Launch form of the last local call: 1 direct, 2 persistent."""
        return self.local_shape()[0]

    def local_shape(self):
        """(launch form, gather depth) of the last local call."""
        m = C.c_int(0)
        check(load().thx_ctx_local_mode(self._h, C.byref(m)), "local_mode")
        return m.value & 15, (m.value >> 4) or 4

    def set_learned_fallback_selection(self, n_rot=0, n_trans=0):
        """Fallback of the learned expectation: rescore flagged particles over
        their proposal's n_rot x n_trans most probable grid poses (0, 0: the
        whole-grid scan)."""
        check(load().thx_ctx_set_learned_fallback_selection(self._h, n_rot, n_trans),
              "set_learned_fallback_selection")

    def set_select_order(self, on=True):
        """Particle order of the learned expectation's selected-pose scorer
        (proposal-mode order when on; index order when off)."""
        check(load().thx_ctx_set_select_order(self._h, int(bool(on))), "set_select_order")

    def set_local_gather(self, ns=0):
        """Samples in flight per gather step of the 16-bit local scorers (4, 8;
        0 = tuned)."""
        check(load().thx_ctx_set_local_gather(self._h, ns), "set_local_gather")

    # ---- particle-filter local search ----
    def pf_set_state(self, quat, trans, k, s01, peakR, transS):
        quat = c32(quat, np.float64).reshape(self.P, -1, 4)
        trans = c32(trans, np.float64).reshape(self.P, -1, 2)
        self.mLR, self.mLT = quat.shape[1], trans.shape[1]
        check(load().thx_pf_set_state(self._h, self.mLR, self.mLT, ptr(quat), ptr(trans),
                                      ptr(c32(k, np.float64)), ptr(c32(s01, np.float64)),
                                      ptr(c32(peakR, np.float64)), ptr(c32(transS, np.float64))),
              "pf_set_state")

    def pf_from_global(self, cls, mLR, mLT, min_std_r, min_std_t, pf_s, transS, seed=1):
        self.mLR, self.mLT = mLR, mLT
        transS = np.broadcast_to(np.asarray(transS, np.float64), (self.P,)).copy()
        check(load().thx_pf_from_global(self._h, cls, mLR, mLT, min_std_r, min_std_t, pf_s,
                                        ptr(transS), seed), "pf_from_global")

    def pf_from_global_classes(self, n_cls, mLR, mLT, min_std_r, min_std_t, pf_s, transS,
                               seed=1):
        """Class draw per particle from the scanned classes 0..n_cls-1, then
        particle states from that class's marginals; returns the classes."""
        self.mLR, self.mLT = mLR, mLT
        transS = np.broadcast_to(np.asarray(transS, np.float64), (self.P,)).copy()
        out = np.zeros(self.P, np.int32)
        check(load().thx_pf_from_global_classes(self._h, n_cls, mLR, mLT, min_std_r, min_std_t,
                                                pf_s, ptr(transS), seed, ptr(out)),
              "pf_from_global_classes")
        return out

    def pf_from_global_given(self, n_cls, mLR, mLT, min_std_r, min_std_t, pf_s, transS, u0R,
                             u0T, u0C=None, pickC=None):
        """pf_from_global(_classes) with the random draws given per particle
        (u0R, u0T resampling offsets; u0C, pickC the class draw); returns the
        classes."""
        self.mLR, self.mLT = mLR, mLT
        P = self.P
        transS = np.broadcast_to(np.asarray(transS, np.float64), (P,)).copy()
        arr = lambda a: None if a is None else c32(np.broadcast_to(a, (P,)), np.float64)
        out = np.zeros(P, np.int32)
        check(load().thx_pf_from_global_given(self._h, n_cls, mLR, mLT, min_std_r, min_std_t,
                                              pf_s, ptr(transS), ptr(arr(u0C)), ptr(arr(pickC)),
                                              ptr(arr(u0R)), ptr(arr(u0T)), ptr(out)),
              "pf_from_global_given")
        return out

    def pf_converge(self, k, s01, after_global=True):
        """Phase at which each particle's search stops under the scripted
        per-phase statistics k [nPh][P][3], s01 [nPh][P][2] (-1: none)."""
        k, s01 = c32(k, np.float64), c32(s01, np.float64)
        nPh, P = k.shape[0], k.shape[1]
        out = np.zeros(P, np.int32)
        check(load().thx_pf_converge(self._h, ptr(k), ptr(s01), nPh, P, int(after_global),
                                     ptr(out)), "pf_converge")
        return out

    def pf_mat4(self, A):
        """(inverse, determinant, top eigenvector) of each 4x4 of A [n][4][4]."""
        A = c32(A, np.float64).reshape(-1, 16)
        n = A.shape[0]
        inv, det, top = np.zeros((n, 4, 4)), np.zeros(n), np.zeros((n, 4))
        check(load().thx_pf_mat4(self._h, ptr(A), n, ptr(inv), ptr(det), ptr(top)), "pf_mat4")
        return inv, det, top

    def pf_refine(self, cls=None, precision=FP32, after_global=False, pf_l=2.0, pf_s=0.5,
                  max_phase=100, seed=1):
        n = C.c_int(0)
        cl = None if cls is None else c32(cls, np.int32)
        check(load().thx_pf_refine(self._h, ptr(cl), precision, int(after_global), pf_l, pf_s,
                                   max_phase, seed, C.byref(n)), "pf_refine")
        return n.value

    def pf_get_state(self):
        P, R, T = self.P, self.mLR, self.mLT
        out = dict(quat=np.zeros((P, R, 4)), trans=np.zeros((P, T, 2)), wR=np.zeros((P, R)),
                   wT=np.zeros((P, T)), k=np.zeros((P, 3)), s01=np.zeros((P, 2)),
                   topR=np.zeros((P, 4)), topT=np.zeros((P, 2)), peakR=np.zeros(P),
                   nPhase=np.zeros(P, np.int32))
        check(load().thx_pf_get_state(self._h, *[ptr(out[k]) for k in
                                                 ("quat", "trans", "wR", "wT", "k", "s01", "topR",
                                                  "topT", "peakR", "nPhase")]), "pf_get_state")
        return out

    def pf_rot_stats(self, quat):
        quat = c32(quat, np.float64)
        P, n = quat.shape[0], quat.shape[1]
        A, mean, k = np.zeros((P, 4, 4)), np.zeros((P, 4)), np.zeros((P, 3))
        w, it = np.zeros((P, n)), np.zeros(P, np.int32)
        check(load().thx_pf_rot_stats(self._h, ptr(quat), n, P, ptr(A), ptr(mean), ptr(k),
                                      ptr(w), ptr(it)), "pf_rot_stats")
        return A, mean, k, w, it

    def pf_trans_stats(self, trans):
        trans = c32(trans, np.float64)
        P, n = trans.shape[0], trans.shape[1]
        m, s, w = np.zeros((P, 2)), np.zeros((P, 2)), np.zeros((P, n))
        check(load().thx_pf_trans_stats(self._h, ptr(trans), n, P, ptr(m), ptr(s), ptr(w)),
              "pf_trans_stats")
        return m, s, w

    def pf_keep_peak(self, u, factor, set_factor):
        u = c32(u, np.float64).copy()
        P, n = u.shape
        f = np.broadcast_to(np.asarray(factor, np.float64), (P,)).copy()
        top = np.zeros(P, np.int32)
        check(load().thx_pf_keep_peak(self._h, ptr(u), n, P, ptr(f), int(set_factor), ptr(top)),
              "pf_keep_peak")
        return u, f, top

    # ---- learned proposal ----
    def expect_learned(self, net, cls=0, n_rot=300, n_trans=25, ess_min=0.0, spread_max=0.0,
                       tail_max=0.0, score_drop=0.0, fallback_precision=BF16):
        """Learned-proposal expectation; returns the number of particles that
        fell back to the whole-grid search (thresholds <= 0 disable a test)."""
        nf = C.c_int(0)
        check(load().thx_ctx_expect_learned(self._h, net._h, cls, n_rot, n_trans, ess_min,
                                            spread_max, tail_max, score_drop,
                                            fallback_precision, C.byref(nf)), "expect_learned")
        return nf.value

    def expect_seeded(self, net, cls=0, mLR=256, mLT=25, precision=BF16, ess_min=20.0,
                      fallback=BF16, min_std_r=0.0, min_std_t=0.0, pf_l=2.0, pf_s=0.5,
                      transS=1.0, max_phase=100, seed=1):
        """Local search seeded by the proposal (see thx_ctx_expect_seeded);
        returns (particles that fell back, phases run)."""
        self.mLR, self.mLT = mLR, mLT
        tS = np.broadcast_to(np.asarray(transS, np.float64), (self.P,)).copy()
        nf, ph = C.c_int(0), C.c_int(0)
        check(load().thx_ctx_expect_seeded(self._h, net._h, cls, mLR, mLT, precision, ess_min,
                                           fallback, min_std_r, min_std_t, pf_l, pf_s, ptr(tS),
                                           max_phase, seed, C.byref(nf), C.byref(ph)),
              "expect_seeded")
        return nf.value, ph.value

    def get_learned(self):
        """(bestR, bestT, score, ess, spread, tail, fell_back) per particle."""
        P = self.P
        bR, bT = np.zeros(P, np.int32), np.zeros(P, np.int32)
        sc, ess = np.zeros(P), np.zeros(P)
        spread, tail = np.zeros(P, np.float32), np.zeros(P, np.float32)
        fell = np.zeros(P, np.int32)
        check(load().thx_ctx_get_learned(self._h, ptr(bR), ptr(bT), ptr(sc), ptr(ess), ptr(spread),
                                         ptr(tail), ptr(fell)), "get_learned")
        return bR, bT, sc, ess, spread, tail, fell

    # ---- particle filter ----
    def pf_resample(self, w, u, nIn, nOut, seed=1, u0=None):
        w, u = c32(w, np.float64), c32(u, np.float64)
        nIn, nOut = c32(nIn, np.int32), c32(nOut, np.int32)
        P = nIn.size
        idx = np.zeros(int(nOut.sum()), np.int32)
        wn = np.zeros(int(nOut.sum()))
        ess = np.zeros(P)
        u0 = None if u0 is None else c32(u0, np.float64)
        check(load().thx_pf_resample(self._h, ptr(w), ptr(u), ptr(nIn), ptr(nOut), P, seed,
                                     ptr(u0), ptr(idx), ptr(wn), ptr(ess)), "pf_resample")
        return idx, wn, ess

    def pf_rot_mean(self, quat, cnt, w=None):
        quat = c32(quat, np.float64)
        cnt = c32(cnt, np.int32)
        w = None if w is None else c32(w, np.float64)
        P = cnt.size
        mean, k = np.zeros((P, 4)), np.zeros((P, 3))
        check(load().thx_pf_rot_mean(self._h, ptr(quat), ptr(w), ptr(cnt), P, ptr(mean), ptr(k)),
              "pf_rot_mean")
        return mean, k

    def pf_perturb(self, quat=None, nR=None, trans=None, nT=None, mean=None, k=None, s01rho=None,
                   pf=1.0, seed=1, phase=0, d=None):
        P = len(nR) if nR is not None else len(nT)
        q = None if quat is None else c32(quat, np.float64).copy()
        t = None if trans is None else c32(trans, np.float64).copy()
        arr = lambda a, dt: None if a is None else c32(a, dt)
        check(load().thx_pf_perturb(self._h, ptr(q), ptr(arr(nR, np.int32)), ptr(t),
                                    ptr(arr(nT, np.int32)), P, ptr(arr(mean, np.float64)),
                                    ptr(arr(k, np.float64)), ptr(arr(s01rho, np.float64)), pf,
                                    seed, phase, ptr(arr(d, np.float64))), "pf_perturb")
        return q, t

    def proposal_sample(self, mix, C, nRot, nTr, seed=1):
        mix = c32(mix, np.float64)
        P = mix.size // (6 * C + 3)
        q = np.zeros((P * nRot, 4))
        dens = np.zeros(P * nRot)
        t = np.zeros((P * nTr, 2))
        check(load().thx_proposal_sample(self._h, ptr(mix), C, P, nRot, nTr, seed, ptr(q),
                                         ptr(dens), ptr(t)), "proposal_sample")
        return q, dens, t

    def last_timing(self):
        buf = C.create_string_buffer(1024)
        ms = np.zeros(16)
        n = C.c_int(0)
        check(load().thx_ctx_last_timing(self._h, buf, 1024, ptr(ms), 16, C.byref(n)),
              "last_timing")
        names = buf.value.decode().split(";") if n.value else []
        return dict(zip(names, ms[: n.value].tolist()))

    def synth_volume(self, cls, idim, pf, nBlob, sigma, rBall, seed):
        cen = np.zeros((nBlob, 3))
        check(load().thx_ctx_synth_volume(self._h, cls, idim, pf, nBlob, sigma, rBall, seed,
                                          ptr(cen)), "synth_volume")
        return cen

    def synth_volume_blobs(self, cls, idim, pf, centres, sigma, amps=None):
        centres = c32(centres, np.float64).reshape(-1, 3)
        a = None if amps is None else c32(amps, np.float64)
        assert a is None or a.size == len(centres)
        check(load().thx_ctx_synth_volume_blobs(self._h, cls, idim, pf, len(centres), sigma,
                                                ptr(centres), ptr(a)), "synth_volume_blobs")

    def synth_particles(self, cls, P, maxShift, dfMin, dfMax, snr, seed, voltage=300e3,
                        Cs=2.7e7, ampContrast=0.07, pixelSize=1.5):
        q = np.zeros((P, 4))
        t = np.zeros((P, 2))
        cp = np.zeros((P, 3))
        check(load().thx_ctx_synth_particles(self._h, cls, P, maxShift, dfMin, dfMax, voltage,
                                             Cs, ampContrast, pixelSize, snr, seed, ptr(q),
                                             ptr(t), ptr(cp)), "synth_particles")
        self.P = P
        return q, t, cp


class Proposal:
    """The learned pose proposal (CNN with hand-written HIP kernels)."""

    def __init__(self, components=8, seed=1, device=0):
        self._h = C.c_void_p()
        self.C = components
        self.NOUT = 6 * components + 3
        check(load().thx_cnn_create(device, components, seed, C.byref(self._h)), "cnn_create")
        n = C.c_size_t(0)
        check(load().thx_cnn_num_params(self._h, C.byref(n)), "cnn_num_params")
        self.nparam = n.value

    def close(self):
        if self._h:
            load().thx_cnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        check(load().thx_cnn_get_params(self._h, ptr(p)), "cnn_get_params")
        return p

    def set_params(self, p):
        p = c32(p)
        assert p.size == self.nparam
        check(load().thx_cnn_set_params(self._h, ptr(p)), "cnn_set_params")

    def forward(self, images):
        images = c32(images)
        B = images.shape[0]
        out = np.zeros((B, self.NOUT), np.float32)
        check(load().thx_cnn_forward(self._h, ptr(images), B, ptr(out)), "cnn_forward")
        return out

    def loss_grad(self, images, q, t, want_grad=True):
        images = c32(images)
        q, t = c32(q, np.float64), c32(t, np.float64)
        loss = C.c_double(0)
        g = np.zeros(self.nparam, np.float32) if want_grad else None
        check(load().thx_cnn_loss_grad(self._h, ptr(images), images.shape[0], ptr(q), ptr(t),
                                       C.byref(loss), ptr(g)), "cnn_loss_grad")
        return loss.value, g

    def train_step(self, images, q, t, lr):
        images = c32(images)
        q, t = c32(q, np.float64), c32(t, np.float64)
        loss = C.c_double(0)
        check(load().thx_cnn_train_step(self._h, ptr(images), images.shape[0], ptr(q), ptr(t),
                                        lr, C.byref(loss)), "cnn_train_step")
        return loss.value

    def train_ctx(self, ctx, q, t, batch=256, epochs=1, lr=1e-3):
        q, t = c32(q, np.float64), c32(t, np.float64)
        loss = C.c_double(0)
        check(load().thx_cnn_train_ctx(self._h, ctx._h, ptr(q), ptr(t), batch, epochs, lr,
                                       C.byref(loss)), "cnn_train_ctx")
        return loss.value

    def set_precision(self, precision):
        """Operand precision of the proposal passes: BF16 (default) or FP32."""
        check(load().thx_cnn_set_precision(self._h, precision), "cnn_set_precision")

    def propose(self, ctx):
        mix = np.zeros((ctx.P, self.NOUT))
        check(load().thx_cnn_propose_ctx(self._h, ctx._h, ptr(mix), None), "cnn_propose_ctx")
        return mix


def ctx_images(ctx, precision=None):
    """64 x 64 images of the context's particles (FP32 transform; BF16: the
    transform the BF16 proposal passes use)."""
    out = np.zeros((ctx.P, 64, 64), np.float32)
    if precision is None:
        check(load().thx_ctx_images(ctx._h, ptr(out)), "ctx_images")
    else:
        check(load().thx_ctx_images_prec(ctx._h, precision, ptr(out)), "ctx_images")
    return out
