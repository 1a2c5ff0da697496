/*
 * thunder_expect.h -- C ABI of the MI355X (gfx950) expectation step.
 *
 * This header is the drop-in boundary for THUNDER's GPU expectation
 * (gpu/interface/Interface.h:16-237, called only from
 * Optimiser::expectationG, src/Optimiser.cpp:1684-3403).  Every entry point
 * takes plain pointers and sizes, returns an int status (THX_OK = 0) and
 * never exits the process; the reference instead prints and calls exit(1)
 * on any CUDA error (gpu/config/Device.cuh.in:28-40).  thx_last_error()
 * returns the message of the calling thread's last failure.
 *
 * Two layers:
 *   1. Reference-shaped entry points (thx_Expect*): same argument meaning,
 *      order and array layouts as the Interface.h functions they replace,
 *      host buffers in and out, synchronous.  Complex is {float re, im}
 *      interleaved, exactly as cuthunder::Complex (reinterpret-cast from
 *      THUNDER's Complex at gpu/interface/Interface.cpp:47-50).
 *   2. A context API (thx_ctx_*) that keeps the volume, pixel set, particle
 *      arrays and pose grids resident in HBM between calls, which is how the
 *      hot loop is driven when the caller owns no host copies.
 *
 * Types: RFLOAT is float (SINGLE_PRECISION=ON, include/Precision.h:65);
 * rotations, translations and prior weights are double
 * (include/Typedef.h:59-64).  Marginals produced by the context API are FP64;
 * the reference-shaped layer converts them to the reference's RFLOAT arrays.
 */
#ifndef THUNDER_EXPECT_H
#define THUNDER_EXPECT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum thx_status {
    THX_OK = 0,
    THX_ERR_ARG = 1,     /* null pointer, bad size or inconsistent shapes */
    THX_ERR_HIP = 2,     /* a HIP runtime call or kernel launch failed   */
    THX_ERR_OOM = 3,     /* device or host allocation failed             */
    THX_ERR_STATE = 4,   /* call order violated (e.g. no volume set)     */
    THX_ERR_RANGE = 5    /* a size exceeds what the kernels support      */
};

/* Arithmetic of the cross-term contraction.  The norm terms, the per-particle
 * maximum and the marginal sums are FP32 inside a tile and FP64 across tiles
 * in every mode. */
enum thx_precision {
    THX_FP32 = 0,   /* plain path: FP32 operands, FP32 MFMA accumulation   */
    THX_BF16 = 1,   /* BF16 operands, FP32 accumulation                    */
    THX_FP16 = 2,   /* FP16 operands with per-shell scales, FP32 accum.    */
    THX_FP8 = 3     /* OCP e4m3 operands with per-shell scales, FP32 accum.*/
};

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* thx_last_error(void);

/* Replaces getAviDevice (gpu/interface/Interface.h:16): number of usable
 * gfx950 devices; the list is 0..n-1. */
int thx_device_count(int* n);

/* ------------------------------------------------------------------------ */
/* 1. Reference-shaped entry points (host buffers, synchronous, device 0).   */
/* ------------------------------------------------------------------------ */

/* Replaces ExpectRotran (gpu/interface/Interface.h:199-208, host side
 * cuthunder::expectRotran gpu/src/cuthunder.cu:1410-1596).
 * traP[nT*npxl] complex out: exp(-2 pi i (iCol tx + iRow ty) / idim);
 * trans[nT*2] in; rot[nR*4] quaternions (w,x,y,z) in;
 * rotMat[nR*9] out, column-major 3x3 per rotation as kernel_getRotMat stores it. */
int thx_ExpectRotran(float* traP, const double* trans, const double* rot,
                     double* rotMat, const int* iCol, const int* iRow,
                     int nR, int nT, int idim, int npxl);

/* Replaces ExpectProject (gpu/interface/Interface.h:210-219, host side
 * gpu/src/cuthunder.cu:1604-1818).  volume: (vdim/2+1)*vdim*vdim complex,
 * x fastest; rotP[nR*npxl] complex out; rotMat[nR*9] column-major;
 * interp must be 1 (LINEAR_INTERP, the only mode the 3-D search uses,
 * include/Model.h:90-93). */
int thx_ExpectProject(const float* volume, float* rotP, const double* rotMat,
                      const int* iCol, const int* iRow, int nR, int pf,
                      int interp, int vdim, int npxl);

/* Replaces ExpectGlobal3D (gpu/interface/Interface.h:221-237, host side
 * gpu/src/cuthunder.cu:1826-2379).  Scores every image against every
 * (rotation, translation) of class kIdx and updates the marginals:
 *   wC[img*nK + kIdx]          = sum_{r,t} exp(dvp - baseL) pR[r] pT[t]
 *   wR[(img*nK + kIdx)*nR + r] += sum_t  exp(dvp - baseL) pT[t]
 *   wT[(img*nK + kIdx)*nT + t] += sum_r  exp(dvp - baseL) pR[r]
 * with dvp = sum_k sigRcp |dat - ctf (traP (.) rotP)|^2 and baseL the running
 * maximum over classes 0..kIdx; when class kIdx raises baseL the entries
 * already stored for earlier classes are rescaled, as kernel_setBaseLine
 * (gpu/src/Kernel.cu:1096-1128) does.  datP/ctfP/sigRcpP are image-major
 * [imgNum*npxl]. */
int thx_ExpectGlobal3D(const float* rotP, const float* traP, const float* datP,
                       const float* ctfP, const float* sigRcpP,
                       float* wC, float* wR, float* wT,
                       const double* pR, const double* pT, float* baseL,
                       int kIdx, int nK, int nR, int nT, int npxl, int imgNum);

/* ------------------------------------------------------------------------ */
/* 2. Context API: data resident in HBM.                                     */
/* ------------------------------------------------------------------------ */

typedef struct thx_ctx thx_ctx;

/* One context per device; owns a stream and all device buffers.  Calls on
 * one context must not overlap; distinct contexts are independent.
 * Replaces the per-device setup of Optimiser::expectationG
 * (src/Optimiser.cpp:1696-1711, 2197-2342). */
int thx_ctx_create(int device, thx_ctx** out);
int thx_ctx_destroy(thx_ctx* ctx);

/* Pixel set of the E-step (Optimiser::allocPreCalIdx, src/Optimiser.cpp:
 * 7942-7992): iCol, iRow, iSig[K]; idim = image size N; pf = padding factor.
 * Replaces ExpectPreidx (gpu/interface/Interface.h:18-23). */
int thx_ctx_set_pixels(thx_ctx* ctx, const int* iCol, const int* iRow,
                       const int* iSig, int K, int idim, int pf);

/* Half-complex padded Fourier volume of class cls, (vdim/2+1)*vdim*vdim
 * complex, x fastest (Projector::projectee3D, include/Projector.h).
 * Replaces ExpectLocalV3D (gpu/interface/Interface.h:45-48) and the volume
 * upload inside ExpectProject. */
int thx_ctx_set_volume(thx_ctx* ctx, int cls, const float* vol, int vdim);

/* Projectee preparation on the GPU: the N^3 half-complex map volFT
 * [(N/2+1)*N*N complex, x fastest] is transformed to real space (scaled by
 * 1/N^3), zero-padded around the origin to V = pf*N, divided by
 * TIK_RL(r / (pf*V)) = j0(pi r / (pf*V))^2 and transformed back; the
 * (V/2+1)*V*V half-complex result becomes class cls's volume (vdim V).
 * Replaces Projector::setProjectee(Volume, nThread) (src/Projector.cpp:123-148)
 * with gridCorrection (src/Projector.cpp:524-606) and the 3-D FFTs of
 * FFT::bw / FFT::fw (src/FFT.cpp:172-230).  msOut (optional, 3 floats):
 * inverse FFT, pad + correction kernel and forward FFT times in ms. */
int thx_ctx_set_projectee(thx_ctx* ctx, int cls, const float* volFT, int N, int pf,
                          float* msOut);
int thx_ctx_get_volume(thx_ctx* ctx, int cls, float* vol);

/* Projection of nR rotations from class cls's resident volume at nPxl given
 * pixels: out[r*nPxl + i] (complex) = the volume's trilinear value at
 * M_r (iCol[i] pf, iRow[i] pf, 0), conjugated mirror for negative x, as
 * Projector::project(Complex*, const dmat33&, iCol, iRow, nPxl, nThread)
 * (src/Projector.cpp:356-374) computes it; rotMat[nR*9] column-major (the
 * storage of an Eigen dmat33 and of kernel_getRotMat).  Synchronous, host
 * buffers. */
int thx_ctx_project(thx_ctx* ctx, int cls, int pf, const double* rotMat, int nR,
                    const int* iCol, const int* iRow, int nPxl, float* out);

/* Particle images to the context's particle data on the GPU: images [P][N][N]
 * real, row-major with the particle centre at pixel (N/2, N/2) as stored in
 * an MRC stack.  Each image's background (centred i^2 + j^2 > bgRadius^2)
 * mean is subtracted and the image divided by the background's sample
 * standard deviation, the image is moved to the centre-at-origin layout, FFT'd
 * (unnormalised, forward) and its pixel set gathered image-major; ctfP and
 * sigRcpP [P*K] are taken as in thx_ctx_set_particles (either may be NULL:
 * thx_ctx_set_ctf and thx_ctx_init_sigma then fill them).  Replaces the image
 * path of Optimiser::initImg (substractBgImg and fwImg,
 * src/Optimiser.cpp:4917-4951, 5003-5013; bgMeanStddev,
 * src/Image/ImageFunctions.cpp:607-620) and the datP packing of
 * Optimiser::allocPreCal (src/Optimiser.cpp:8005-8016).  With softMask
 * nonzero the packed data are those of the soft-masked images, as the
 * expectation packs them (allocPreCal(mask = true), src/Optimiser.cpp:637,
 * 1457; maskImg / reMaskImg with zeroMask, src/Optimiser.cpp:4953-4985,
 * 6053-6105): each normalised image times the mask of radius bgRadius with
 * a raised-cosine edge of 6 pixels (softMask, src/Functions/Mask.cpp:333-350;
 * EDGE_WIDTH_RL, include/Macro.h:82); with softMask zero, those of the
 * unmasked images (allocPreCal(mask = false)).  The unmasked spectra stay
 * resident for thx_ctx_init_sigma either way.  msOut (optional, 3 floats):
 * normalisation, FFT and packing times in ms (the host upload not included). */
int thx_ctx_set_particle_images(thx_ctx* ctx, const float* images, int P, int N, float bgRadius,
                                int softMask, const float* ctfP, const float* sigRcpP,
                                float* msOut);

/* Initial noise spectrum from the images of the last
 * thx_ctx_set_particle_images call (their normalised, unmasked spectra):
 * sigma^2(u) = (average over images of the power spectrum in shell u
 * - (ring average of Re + Im of the average image)^2) / 2 and
 * sigRcp(u) = -0.5 / sigma^2(u), for shells u = 0 .. nShell - 1 with
 * nShell <= N/2 - 1, shell u holding the pixels with rint(|(i, j)|) == u.
 * Replaces Optimiser::initSigma (src/Optimiser.cpp:5111-5209; powerSpectrum and
 * ringAverage, src/Functions/Spectrum.cpp:47-62, 161-190) for one group;
 * nShell = maxR() (src/Optimiser.cpp:4384-4387) gives its full output.  With apply
 * nonzero the context's per-particle sigRcpP becomes sigRcp(iSig[k]) for every
 * particle, as allocPreCal fills it (src/Optimiser.cpp:8017-8022).  sigOut,
 * sigRcpOut (optional, nShell doubles each); msOut (optional, 1 float):
 * kernel time in ms. */
int thx_ctx_init_sigma(thx_ctx* ctx, int nShell, int apply, double* sigOut, double* sigRcpOut,
                       float* msOut);

/* Optimiser::statImg's noise deviation and normaliseImg's scale for the last
 * thx_ctx_set_particle_images call (src/Optimiser.cpp:4796-4895, 4987-5001):
 * stdN = the mean over its P images of bgStddev(0, image, bgRadius), GSL's
 * standard deviation about 0 with n - 1, of the background-subtracted
 * images; scale = 1 / stdN rounded to float.  The packed data and the initial
 * sigma carry that scale, as normaliseImg applies it to every image before
 * fwImg.  Either output may be NULL. */
int thx_ctx_get_image_stats(thx_ctx* ctx, double* stdN, float* scale);

/* The noise spectrum refreshed from each particle's best pose after an
 * expectation (Optimiser::allReduceSigma, src/Optimiser.cpp:6355-6666, 3-D,
 * best pose only, all shells to maxR, masked and unmasked residuals mixed,
 * recentred images).  Particle p's pose: class cls[p] (NULL: 0), quaternion
 * quat[p*4], translation trans[p*2] (pixels) and recentring offset
 * offset[p*2] (NULL: 0); group[p] in 0 .. nGroup-1 (NULL: one group).  The
 * model CTF . shift . slice is formed on the half-spectrum pixels with
 * |k| < r (the projector's maxRadius, <= N/2 - 1) from the class volumes and
 * the particles' CTF attributes (thx_ctx_set_ctf), and compared with the
 * masked and unmasked spectra of the last thx_ctx_set_particle_images call
 * (with their normaliseImg scale).  Per group and shell u < nShell
 * (<= N/2 - 1): sigM, sigN = mean over the group's particles of half the
 * mean residual power of the shell, masked and unmasked; svd = mean of
 * sqrt(model power / masked data power); sigma^2 = ratio sigM + (1 - ratio)
 * alpha sigN with ratio = min(1, svd) and alpha = sqrt(pi) maskRadius /
 * (N pixelSize) (A); sigRcp = -0.5 / sigma^2.  sigOut, sigRcpOut
 * [nGroup*nShell] (optional); apply != 0 writes sigRcp(group, iSig[k]) into
 * every particle's sigRcpP as allocPreCal does.  msOut (optional, 1 float):
 * kernel time. */
int thx_ctx_refresh_sigma(thx_ctx* ctx, const int* cls, const double* quat, const double* trans,
                          const double* offset, const int* group, int nGroup, int r, int nShell,
                          float maskRadius, float pixelSize, int apply, double* sigOut,
                          double* sigRcpOut, float* msOut);

/* The intensity scale per group from the best poses (Optimiser::refreshScale,
 * src/Optimiser.cpp:5715-6023, and correctScale, :5069-5109; 3-D; the
 * unmasked images; one ratio of sums per group).  With coord != 0 particle
 * p's model is class cls[p]'s slice at quat[p*4], shifted by trans[p*2] -
 * offset[p*2] (offset NULL: 0); with coord == 0 (the first iteration) it is
 * class 0's unshifted slice at quat[p*4] (NULL: identity) and only the zero
 * frequency counts (rS = 1, rL = 0), as the reference's
 * OPTIMISER_REFRESH_SCALE_ZERO_FREQ_NO_COORD sets it.  Over the stored pixels
 * with rL^2 <= i^2 + j^2 < rS^2 and rL <= rint|k| < rS (rS clamped to r, the
 * projector's radius <= N/2 - 1), with the plain complex products the
 * reference writes (scaleDataVSPrior, :10021-10063):
 *   sXA = sum Re(X M) c,  sAA = sum Re(M M) c^2,
 * X the unmasked data, M the model slice, c the CTF from the attributes.
 * scale[g] = sXA / sAA over group g's particles (group NULL: over all
 * particles, every group the same); then, with med the median of the nGroup
 * scales, a scale above |5 med| in magnitude becomes 5 med and one below
 * |med / 5| becomes med / 5.  A negative mean is THX_ERR_RANGE.  apply: 0
 * nothing; 1 (correctScale, init = false) divides every particle's spectra
 * and packed data by scale[group[p]] and multiplies its sigRcpP by the
 * square (sigma^2 / scale^2); 2 (init = true) multiplies every class volume
 * by scale[group[0]].  scaleOut [nGroup] (optional); msOut: kernel time. */
int thx_ctx_refresh_scale(thx_ctx* ctx, int coord, const int* cls, const double* quat,
                          const double* trans, const double* offset, const int* group,
                          int nGroup, int r, int rS, float rL, float pixelSize, int apply,
                          double* scaleOut, float* msOut);

/* Noise normalisation from the best poses (Optimiser::normCorrection,
 * src/Optimiser.cpp:6161-6351; 3-D, no CTF search, the masked images
 * (OPTIMISER_NORM_MASK) with the translation trans[p*2] not recentred):
 * norm[p] = sum over the stored pixels with rL^2 <= i^2 + j^2 < rNorm^2
 * (rNorm clamped to r) of |X_masked - c M|^2, M class cls[p]'s slice at
 * quat[p*4] shifted by trans[p*2], c the CTF; m = the median of the norms
 * (GSL's 0.5 quantile, rounded to float).  apply != 0 multiplies every
 * particle's spectra (masked and unmasked) and packed data by
 * sqrt(m / norm[p]).  normOut [P], medianOut [1] (optional); msOut: kernel
 * time. */
int thx_ctx_norm_correction(thx_ctx* ctx, const int* cls, const double* quat,
                            const double* trans, int r, float rL, float rNorm, float pixelSize,
                            int apply, double* normOut, double* medianOut, float* msOut);

/* Particle arrays, image-major [P*K] as allocPreCal(pixelMajor = false)
 * packs them (src/Optimiser.cpp:7994-8121).  Replaces ExpectLocalIn +
 * ExpectLocalP (gpu/interface/Interface.h:31-38, 50-62) for all particles at once.
 * ctfP and sigRcpP may be NULL: then thx_ctx_set_ctf and thx_ctx_init_sigma
 * (apply) fill them on the device; scoring before they are set is
 * THX_ERR_STATE. */
int thx_ctx_set_particles(thx_ctx* ctx, const float* datP, const float* ctfP,
                          const float* sigRcpP, int P);

/* The particles' CTF at the pixel set, evaluated on the device from their CTF
 * attributes: ctfAttr[P*7] = (voltage V, defocusU A, defocusV A, defocusTheta
 * rad, Cs A, amplitudeContrast, phaseShift rad) per particle, the field order
 * of CTFAttr (include/Database.h:108-124), so a std::vector<CTFAttr> passes
 * as is; pixelSize in A.  ctf = -sqrt(1 - A^2) sin ki + A cos ki with
 * ki = pi lambda defocus(angle) u^2 + pi/2 Cs lambda^3 u^4 - phaseShift
 * (CTF(RFLOAT*, ...), src/CTF.cpp:109-145), as allocPreCal packs ctfP with
 * OPTIMISER_CTF_ON_THE_FLY (src/Optimiser.cpp:8033-8057) after initCTF
 * (:5027-5068).  P must be the context's particle count. */
int thx_ctx_set_ctf(thx_ctx* ctx, const float* ctfAttr, int P, float pixelSize);
int thx_ctx_get_particles(thx_ctx* ctx, float* datP, float* ctfP,
                          float* sigRcpP);

/* Global pose grid: quaternions (w,x,y,z) [R*4], translations in pixels
 * [T*2] and their prior weights pR[R], pT[T] (Particle::reset,
 * src/Particle.cpp:87-169, and the par.t / par.rot calls of
 * src/Optimiser.cpp:711-725, 768-782). */
int thx_ctx_set_grid(thx_ctx* ctx, const double* quat, int R,
                     const double* trans, int T, const double* pR,
                     const double* pT);

/* Global scan of class cls at the given precision (the doGlobalSearch hot
 * loop, src/Optimiser.cpp:757-915, and ExpectRotran + ExpectProject +
 * ExpectGlobal3D on the GPU path).  Results stay on the device until
 * thx_ctx_get_global is called.  Enqueued on the context stream; returns
 * once the work is queued. */
int thx_ctx_score_global(thx_ctx* ctx, int cls, int precision);

/* Copies the marginals of the last scan of every class merged over classes
 * with one baseline: wC[P*nK], wR[(P*nK)*R], wT[(P*nK)*T], baseL[P] (FP64),
 * the same layouts as thx_ExpectGlobal3D.  Synchronises the stream. */
int thx_ctx_get_global(thx_ctx* ctx, int nK, double* wC, double* wR,
                       double* wT, double* baseL);

/* Per-particle best pose of the last scan of class cls (argmax over r,t):
 * bestR[P], bestT[P] (indices) and bestScore[P] = max dvp. */
int thx_ctx_get_best(thx_ctx* ctx, int cls, int* bestR, int* bestT,
                     double* bestScore);

/* Scores of particle p against every (r, t) of the grid, FP32, for
 * checking: out[R*T] (r slowest).  Uses the plain FP32 path. */
int thx_ctx_score_one(thx_ctx* ctx, int cls, int p, float* out);

/* One local-search phase for every particle of the context, batched and
 * ragged (Optimiser::updateSP, src/Optimiser.cpp:1139-1437; on the GPU path
 * ExpectLocalRTD + ExpectLocalPreI3D + ExpectLocalM, gpu/interface/
 * Interface.h:80-87, 107-123, 125-139, which handle one particle per call).  Particle p scores
 * its own nR[p] rotations (quaternions, concatenated over particles) and nT[p]
 * <= 128 translations (pixels, concatenated) with prior weights pR, pT (same
 * concatenation), against the volume of class cls[p] (cls may be NULL: class
 * 0).  Slices are projected on the fly; no R x K intermediate is stored.
 * precision: THX_FP32, THX_BF16 or THX_FP16 (per-shell operand scales).
 * Marginals as in the global scan:
 *   wC[p] = sum w pR pT, wR[r] = sum_t w pT, wT[t] = sum_r w pR,
 *   w = exp(dvp - baseL[p]), baseL[p] = max dvp.
 * nR[p] = 0 skips particle p (nT[p] may then be 0): wC[p] = 0,
 * baseL[p] = -inf, no best pose (-1); at least one particle must be scored. */
int thx_ctx_score_local(thx_ctx* ctx, const int* cls, const double* quat,
                        const int* nR, const double* trans, const int* nT,
                        const double* pR, const double* pT, int precision);

/* Results of the last local phase: wC[P], wR[sum nR], wT[sum nT], baseL[P],
 * bestR[P] / bestT[P] (indices within the particle's own lists).  Any
 * output may be NULL.  Synchronises the stream. */
int thx_ctx_get_local(thx_ctx* ctx, double* wC, double* wR, double* wT,
                      double* baseL, int* bestR, int* bestT);

/* Launch shape of the local scorer (no reference counterpart): mode 0 =
 * autotuned (the candidate shapes are timed on the first batch of each size
 * class and the fastest is kept), 1 = one block per work item (particle x
 * 128 rotations), 2 = a persistent grid sized to the resident capacity
 * pulling work items from a counter.  graph != 0 captures the local call
 * (operand prep, scorers, merge) and the particle filter's phases in
 * hipGraphs, replayed while the batch shape and the context's buffers stay
 * the same. */
int thx_ctx_set_local_options(thx_ctx* ctx, int mode, int graph);
/* Gather depth of the 16-bit local scorers (no reference counterpart): the
 * trilinear samples per step whose corner loads are all in flight, 4 or 8;
 * 0 (the default) lets the tuner time both. */
int thx_ctx_set_local_gather(thx_ctx* ctx, int ns);
/* Launch shape used by the last local call: the form (1 or 2) plus 16 x the
 * gather depth. */
int thx_ctx_local_mode(thx_ctx* ctx, int* mode);

/* Waits for all work queued on the context stream. */
int thx_ctx_sync(thx_ctx* ctx);

/* Timing of the last thx_ctx_score_global: per-kernel average duration in
 * milliseconds measured with HIP events on the context stream. names is a
 * ';'-separated list written into buf (cap bytes); ms[i] in the same order. */
int thx_ctx_last_timing(thx_ctx* ctx, char* names, int cap, double* ms,
                        int maxn, int* n);

/* ------------------------------------------------------------------------ */
/* Particle filter (host buffers, ragged per-particle counts; FP64).         */
/* ------------------------------------------------------------------------ */

/* Systematic resampling for P particles (Particle::resample(n, pt),
 * src/Particle.cpp:1275-1462, one sample type): particle p has nIn[p] <= 8192
 * samples with prior w and likelihood u (concatenated); W = w (.) u
 * normalised, FP64 cumulative sum, nOut[p] systematic draws u0 + j / nOut
 * with u0 uniform in [0, 1/nOut) (u0 != NULL: given per particle), new
 * weight 1 / u (PARTICLE_PRIOR_ONE) normalised.  idx (within the particle),
 * wNew: concatenated over nOut.  ess[p] = 1 / sum W^2 (optional; the
 * commented-out Particle::neff, src/Particle.cpp:1647-1651).  The random
 * shuffle before resampling (src/Particle.cpp:2173-2271) is not applied. */
int thx_pf_resample(thx_ctx* ctx, const double* w, const double* u,
                    const int* nIn, const int* nOut, int P, uint64_t seed,
                    const double* u0, int* idx, double* wNew, double* ess);

/* No counterpart in the reference: thx_pf_resample for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * Every array is a device pointer on the context's device: w, u [sum nIn],
 * nIn, nOut [P], u0 [P] or NULL, idx, wNew [sum nOut], ess [P] or NULL, with
 * the same meaning and results as thx_pf_resample.  Calls no allocation
 * routine: the per-particle offsets go into a static device array of
 * THX_NOALLOC_MAX_P particles (2 MiB, held by each GPU it runs on until the
 * process ends); larger P is THX_ERR_RANGE.  Counts are read on the device:
 * a particle whose nIn or nOut lies outside [1, 8192] gets ess = -1 and no
 * outputs (its slots are counted as min(count, 8192)).  Enqueued on the
 * context stream; returns once queued (thx_ctx_sync waits).  Calls that
 * overlap on one GPU share that array and race. */
#define THX_NOALLOC_MAX_P 131072
int thx_pf_resample_noalloc(thx_ctx* ctx, const double* w, const double* u,
                            const int* nIn, const int* nOut, int P, uint64_t seed,
                            const double* u0, int* idx, double* wNew, double* ess);

/* Weighted mean rotation per particle: principal eigenvector of the scatter
 * sum w q q^T (the ACG mean inferACG estimates, src/Geometry/
 * DirectionalStat.cpp:87-244; sign chosen with w >= 0) and the eigenvalue
 * ratios k[3] (Particle::calVari, src/Particle.cpp:1002-1126).  w may be NULL. */
int thx_pf_rot_mean(thx_ctx* ctx, const double* quat, const double* w,
                    const int* cnt, int P, double* mean, double* k);

/* Perturbation (Particle::perturb, src/Particle.cpp:1133-1273) in place:
 * rotations q <- mean d conj(mean) q with d ~ ACG(1, pf^2 k1, pf^2 k2,
 * pf^2 k3) (or the given d[4 * sum nR] when d != NULL); translations
 * t += pf (x, y), (x, y) bivariate Gaussian (s0, s1, rho) per particle.
 * Either part may be skipped by passing NULL quat or NULL trans. */
int thx_pf_perturb(thx_ctx* ctx, double* quat, const int* nR, double* trans,
                   const int* nT, int P, const double* mean, const double* k,
                   const double* s01rho, double pf, uint64_t seed, int phase,
                   const double* d);

/* Samples of the learned proposal: per particle, mix holds C components
 * (m0 m1 m2 m3 k pi) and then (mu_x mu_y sigma) of the shift Gaussian.
 * Writes nRot quaternions quat[P*nRot*4] drawn from the ACG mixture, their
 * mixture density qdens[P*nRot] and nTr shifts trans[P*nTr*2]. */
int thx_proposal_sample(thx_ctx* ctx, const double* mix, int C, int P,
                        int nRot, int nTr, uint64_t seed, double* quat,
                        double* qdens, double* trans);

/* ------------------------------------------------------------------------ */
/* Local search with the particle filter on the GPU (state resident in HBM). */
/* Uniform counts: mLR rotations and mLT <= 128 translations per particle    */
/* (_para.mLR / _para.mLT).  The state is FP64.                              */
/* ------------------------------------------------------------------------ */

/* Particle states of the context's P particles: quat[P*mLR*4], trans[P*mLT*2]
 * (pixels), concentrations k[P*3], shift deviations s01[P*2], the rotation
 * peak factor peakR[P] and the translation search scale transS[P]
 * (Particle::load / the state Particle keeps, src/Particle.cpp:401-554). */
int thx_pf_set_state(thx_ctx* ctx, int mLR, int mLT, const double* quat,
                     const double* trans, const double* k, const double* s01,
                     const double* peakR, const double* transS);

/* Particle states from the last global scan of class cls (src/Optimiser.cpp:
 * 926-1084): the rotation marginal gets its peak factor and half-height peak
 * (setPeakFactor + keepHalfHeightPeak), mLR rotations and mLT shifts are
 * resampled from the grid (prior = the grid's), calVari gives k and s, and
 * the scanning minima apply (k >= (minStdR / pfS)^2, s >= minStdT / pfS). */
int thx_pf_from_global(thx_ctx* ctx, int cls, int mLR, int mLT, double minStdR,
                       double minStdT, double pfS, const double* transS,
                       uint64_t seed);

/* As thx_pf_from_global over the classes 0..nCls-1 scanned last
 * (src/Optimiser.cpp:926-966): per particle the class marginals on a common
 * baseline, u_k = wC_k exp(base_k - max base), keep their half-height peak
 * with the fixed factor 1 - 1e-2 (setPeakFactor / keepHalfHeightPeak(PAR_C),
 * src/Particle.cpp:1879-1886, 1935-1944), are resampled to k systematic
 * draws and one draw is taken uniformly (resample(k, PAR_C) + rand(cls),
 * src/Particle.cpp:1275-1462, 2080-2089); the particle's rotations and
 * shifts then come from that class's marginals.  clsOut[P]: the classes. */
int thx_pf_from_global_classes(thx_ctx* ctx, int nCls, int mLR, int mLT, double minStdR,
                               double minStdT, double pfS, const double* transS,
                               uint64_t seed, int* clsOut);

/* As thx_pf_from_global / thx_pf_from_global_classes with the random draws
 * given instead of taken from a seed (checking, and callers that keep their
 * own random streams): u0R[P] in [0, 1/mLR) and u0T[P] in [0, 1/mLT) are the
 * systematic-resampling offsets of each particle's rotations and shifts
 * (gsl_ran_flat in Particle::resample, src/Particle.cpp:1324-1366); with
 * nCls > 1, u0C[P] in [0, 1/nCls) is the class resampling offset and
 * pickC[P] in [0, 1) picks the draw floor(pickC nCls) (rand(cls),
 * src/Particle.cpp:2080-2089).  nCls == 1 uses class 0 (u0C, pickC may be
 * NULL).  clsOut[P] (optional with nCls == 1): the classes. */
int thx_pf_from_global_given(thx_ctx* ctx, int nCls, int mLR, int mLT, double minStdR,
                             double minStdT, double pfS, const double* transS,
                             const double* u0C, const double* pickC, const double* u0R,
                             const double* u0T, int* clsOut);

/* Local search phases until every particle converges or maxPhase (the
 * expectation loop, src/Optimiser.cpp:1475-1660, with updateSP :1139-1437):
 * per phase, perturbation (pfL at phase 0, else pfS) with balanced weights,
 * scoring of every (rotation, shift) sample against the particle's class
 * (cls[P] or NULL: class 0) at precision THX_FP32, THX_BF16 or THX_FP16
 * (per-shell operand scales bounded from the class volumes), half-height
 * peak of the rotation marginal, top samples, calVari, resampling, and the
 * variance test (from phase 10 after a global scan, afterGlobal != 0, where
 * phases start at 1; else from phase 3).  *phasesRun: phases executed. */
int thx_pf_refine(thx_ctx* ctx, const int* cls, int precision, int afterGlobal,
                  double pfL, double pfS, int maxPhase, uint64_t seed,
                  int* phasesRun);

/* Current state (any output may be NULL): samples, weights, k, s01, the top
 * rotation / shift of the last phase, peak factors and the phase at which
 * each particle converged (-1: not converged). */
int thx_pf_get_state(thx_ctx* ctx, double* quat, double* trans, double* wR,
                     double* wT, double* k, double* s01, double* topR,
                     double* topT, double* peakR, int* nPhase);

/* Stage kernels on host buffers (checking): per set of n rotations, the
 * inferACG matrix A[16], mean[4], calVari k[3], balance weights[n] and the
 * fixed-point iteration count (src/Geometry/DirectionalStat.cpp:87-244,
 * src/Particle.cpp:1002-1081, 2301-2311). */
int thx_pf_rot_stats(thx_ctx* ctx, const double* quat, int n, int P, double* A,
                     double* mean, double* k, double* wBal, int* iters);
/* Per set of n >= 2 shifts: means[2], n - 1 deviations[2], balance weights[n]
 * (src/Particle.cpp:1083-1097, 2313-2344). */
int thx_pf_trans_stats(thx_ctx* ctx, const double* trans, int n, int P,
                       double* m, double* s01, double* wBal);
/* The convergence test of the local search (src/Optimiser.cpp:1480-1612,
 * OPTIMISER_COMPRESS_CRITERIA) on scripted statistics: after phase
 * phase0 + i (phase0 = 1 with afterGlobal, else 0) particle p has
 * concentrations k[(i P + p) * 3 + a] and shift deviations
 * s01[(i P + p) * 2 + a]; nPhase[p] receives the phase that ended its search
 * (-1: none within nPh phases). */
int thx_pf_converge(thx_ctx* ctx, const double* k, const double* s01, int nPh, int P,
                    int afterGlobal, int* nPhase);
/* The 4x4 helpers of the inferACG stages on n matrices A[n*16] (row-major):
 * inv[n*16] and det[n] (inverse zero for a singular matrix), top[n*4] the
 * unit eigenvector of the largest eigenvalue of a symmetric A, its
 * largest-magnitude entry positive (src/Geometry/DirectionalStat.cpp:87-244). */
int thx_pf_mat4(thx_ctx* ctx, const double* A, int n, double* inv, double* det, double* top);
/* setPeakFactor (setFactor != 0) + keepHalfHeightPeak of P sets of n values
 * (in place), factor[P] in/out, top[P] = first maximum (src/Particle.cpp:
 * 988-1000, 1864-1954). */
int thx_pf_keep_peak(thx_ctx* ctx, double* u, int n, int P, double* factor,
                     int setFactor, int* top);

/* ------------------------------------------------------------------------ */
/* Learned pose proposal (no counterpart in the reference, which always     */
/* scans a uniform grid first): a ~0.9M-parameter CNN on 64 x 64 phase-     */
/* flipped particle images predicting a C-component ACG mixture over        */
/* rotations and a Gaussian over shifts.  Training, inference and gradients */
/* run in hand-written HIP kernels.                                          */
/* ------------------------------------------------------------------------ */

typedef struct thx_cnn thx_cnn;

int thx_cnn_create(int device, int components, uint64_t seed, thx_cnn** out);
int thx_cnn_destroy(thx_cnn* net);
int thx_cnn_num_params(thx_cnn* net, size_t* count);
/* Flattened parameters: conv1..4 (weights [co][ci][ky][kx], biases), fc1
 * [384][2048] + bias, fc2 [6C+3][384] + bias. */
int thx_cnn_get_params(thx_cnn* net, float* p);
int thx_cnn_set_params(thx_cnn* net, const float* p);
/* out[B][6C+3]: raw network outputs for images [B][64][64], B <= 512. */
int thx_cnn_forward(thx_cnn* net, const float* images, int B, float* out);
/* Mean mixture negative log-likelihood of true quaternions q[B][4] and shifts
 * t[B][2]; grad (optional) receives d loss / d params. */
int thx_cnn_loss_grad(thx_cnn* net, const float* images, int B, const double* q,
                      const double* t, double* loss, float* grad);
/* One Adam step (lr) on a batch; loss is the batch loss before the step. */
int thx_cnn_train_step(thx_cnn* net, const float* images, int B, const double* q,
                       const double* t, float lr, double* loss);
/* 64 x 64 images of the context's particles (device; optional host copy). */
int thx_ctx_images(thx_ctx* ctx, float* host);
/* The same images made as the proposal passes make them (no reference
 * counterpart): precision THX_FP32 (as thx_ctx_images) or THX_BF16 (the
 * transform on BF16 operands the BF16 proposal passes use). */
int thx_ctx_images_prec(thx_ctx* ctx, int precision, float* host);
/* Train on the context's particles with known poses q[P][4], t[P][2]. */
int thx_cnn_train_ctx(thx_cnn* net, thx_ctx* ctx, const double* q, const double* t,
                      int B, int epochs, float lr, double* lastLoss);
/* Mixture parameters per particle of the context, mix[P][6C+3]: C x (m0 m1
 * m2 m3 k pi) then (mu_x mu_y sigma); host and/or device destination. */
int thx_cnn_propose_ctx(thx_cnn* net, thx_ctx* ctx, double* mixHost, double* mixDev);
int thx_cnn_components(thx_cnn* net, int* C);
/* Operand precision of the proposal passes (thx_cnn_propose_ctx and the
 * learned expectation): THX_BF16 (default; FP32 accumulation) or THX_FP32.
 * Training and thx_cnn_forward always use FP32.  No counterpart in the
 * reference (it has no learned proposal). */
int thx_cnn_set_precision(thx_cnn* net, int precision);

/* Expectation with the learned proposal over the context's grid: the grid's
 * slices are projected once (BF16 table + FP32 |P|^2), the network's mixture
 * ranks the grid for each particle, the nSelR most probable rotations and
 * nSelT (<= 128) most probable translations are scored (BF16 operands, FP32
 * accumulation, FP64 merge).  A particle falls back to the uniform search of
 * the whole grid (the tiled scan at fallbackPrecision) when any enabled test
 * fires (a threshold <= 0 disables its test):
 *   ESS of the importance weights w = wR / q over its scored rotations,
 *     (sum w)^2 / sum w^2, below essMin;
 *   spread of the proposal, the effective number of grid rotations under q,
 *     (sum q)^2 / sum q^2, above spreadMax;
 *   tail depth of its best rotation, max log q - log q(best), above tailMax;
 *   best log-likelihood below the batch median by more than scoreDrop.
 * *nFallback receives the number of particles that fell back. */
int thx_ctx_expect_learned(thx_ctx* ctx, thx_cnn* net, int cls, int nSelR,
                           int nSelT, double essMin, double spreadMax,
                           double tailMax, double scoreDrop,
                           int fallbackPrecision, int* nFallback);
/* Block order of the learned expectation's selected-pose scorer (no
 * reference counterpart): on = 1 (the default) sorts the particles by a
 * Morton code of their proposal's dominant mode and deals them so that each
 * XCD scores a contiguous run of that order (overlapping selections share its
 * L2); 0 takes them in index order.  The results are the same either way. */
int thx_ctx_set_select_order(thx_ctx* ctx, int on);
/* Fallback of the learned expectation (no reference counterpart): with
 * nSelR > 0 a flagged particle is rescored over the nSelR (<= 16384) rotations
 * and nSelT (<= 128) translations of the grid its proposal ranks highest, by
 * the selected-pose scorer, instead of the whole-grid scan (fallbackPrecision
 * is then unused); 0, 0 (the default) restores the whole-grid scan.  The ESS
 * and tail depth that thx_ctx_get_learned returns stay those of the first
 * selection. */
int thx_ctx_set_learned_fallback_selection(thx_ctx* ctx, int nSelR, int nSelT);
/* Local-search expectation seeded by the learned proposal, in place of
 * Particle::reset / load (src/Particle.cpp:87-169, 401-554): per particle,
 * mLR rotations and mLT translations are drawn from the network's mixture
 * (as thx_proposal_sample), their ACG concentrations, shift deviations and
 * balance weights taken (calVari, balanceWeight), and every particle's
 * samples scored once by the local scorer (precision; projection fused); the
 * peak factor, half-height peak, resampling by the scores, calVari and the
 * scanning minima (minStdR, minStdT, pfS) then give the particle-filter
 * state.  Where the ESS of a particle's rotation weights, (sum w)^2 /
 * sum w^2, is below essMin (<= 0: never), the uniform global search is the
 * fallback: the whole grid of the context scanned at fallbackPrecision and
 * the state from that scan (as thx_pf_from_global).  Then the local phases
 * of thx_pf_refine for every particle (after-global phase rules, pfL, pfS,
 * maxPhase, seed + 1); transS[P] is the shift search scale per particle.
 * *nFallback and *phasesRun (optional) receive the particles that fell back
 * and the phases run; the state is read with thx_pf_get_state.  No
 * counterpart in the reference (it has no learned proposal). */
int thx_ctx_expect_seeded(thx_ctx* ctx, thx_cnn* net, int cls, int mLR, int mLT,
                          int precision, double essMin, int fallbackPrecision,
                          double minStdR, double minStdT, double pfL, double pfS,
                          const double* transS, int maxPhase, uint64_t seed,
                          int* nFallback, int* phasesRun);
/* Per particle: best grid rotation / translation indices, best
 * log-likelihood, ESS, spread, tail depth and whether it fell back (any
 * output may be NULL). */
int thx_ctx_get_learned(thx_ctx* ctx, int* bestR, int* bestT, double* score,
                        double* ess, float* spread, float* tail,
                        int* fellBack);

/* ------------------------------------------------------------------------ */
/* Synthetic workloads (seeded; Philox4x32-10 counter streams).              */
/* ------------------------------------------------------------------------ */

/* Volume of nBlob isotropic Gaussian blobs (width sigma voxels, amplitude 1,
 * centres uniform in a ball of radius rBall about the box centre) of an
 * idim^3 box, generated directly as its Fourier transform on the pf-padded
 * half-complex grid (vdim = pf*idim).  centresOut[nBlob*3] (optional). */
int thx_ctx_synth_volume(thx_ctx* ctx, int cls, int idim, int pf, int nBlob,
                         double sigma, double rBall, uint64_t seed,
                         double* centresOut);
/* The same from given blob centres centres[nBlob*3] (voxels from the box
 * centre) and amplitudes amps[nBlob] (NULL: all 1), e.g. a map perturbed as
 * one refinement iteration would change it. */
int thx_ctx_synth_volume_blobs(thx_ctx* ctx, int cls, int idim, int pf, int nBlob,
                               double sigma, const double* centres,
                               const double* amps);

/* Particle images X = CTF (.) S_t (.) P_r + complex white noise at the
 * given real-space SNR, at the context's pixel set, from class cls.
 * True quaternion uniform on S^3, shifts uniform in [-maxShift, maxShift]^2,
 * defocus U, V uniform in [dfMin, dfMax] Angstrom, astigmatism angle
 * uniform in [0, pi).  CTF: voltage V, Cs Angstrom, amplitude contrast,
 * pixelSize Angstrom/px.  sigRcp = -1 / (N^2 sigma_n^2) on every pixel.
 * trueQuat[P*4], trueShift[P*2], ctfParams[P*3] (dU, dV, theta) optional. */
int thx_ctx_synth_particles(thx_ctx* ctx, int cls, int P, double maxShift,
                            double dfMin, double dfMax, double voltage,
                            double Cs, double ampContrast, double pixelSize,
                            double snr, uint64_t seed, double* trueQuat,
                            double* trueShift, double* ctfParams);

/* ------------------------------------------------------------------------ */
/* Consumer side: reconstruction insert.                                     */
/* ------------------------------------------------------------------------ */

/* The insert loop of Optimiser::reconstructRef (src/Optimiser.cpp:7000-7190)
 * for the context's particles (from thx_ctx_set_particle_images, CTF from
 * thx_ctx_set_ctf): particle p's m samples s = p m .. p m + m - 1 with class
 * cls[s] (NULL: 0), quaternion quat[s*4] and shift trans[s*2] (pixels), its
 * recentring offset offset[p*2] (NULL: 0) and weight w[p] (NULL: 1), each
 * inserted with weight w / m (the _par[l].rand draws of mReco samples).
 * Each sample's unmasked data at the pixel set of radius rU
 * (allocPreCalIdx(rU, 0), allocPreCal(mask = false)), shifted by
 * -(t - offset) and times the particle's CTF, is added into the class's F
 * volume and CTF^2 into its T volume by trilinear weights at
 * R (iCol pf, iRow pf, 0) (Reconstructor::insertP, src/Reconstructor.cpp:
 * 782-860; Volume::addFT, src/Image/Volume.cpp:334-368, 555-700), and
 * -R (t - offset, 0) is summed per class (insertDir).  Accumulators: FP64 on
 * the (V/2+1)*V*V half grid of V = pf N per class 0 .. nCls-1, zeroed when
 * clear != 0 (or the size changed), else added to.  msOut (optional): kernel
 * time in ms. */
int thx_ctx_insert(thx_ctx* ctx, const int* cls, const double* quat, const double* trans,
                   const double* offset, int m, const double* w, int rU, int nCls, int clear,
                   float pixelSize, float* msOut);
/* Accumulated F [(V/2+1)*V*V complex, x fastest] and T [(V/2+1)*V*V] of class
 * cls (FP64), the insertDir sum dir[3] and the number of samples inserted
 * (any output may be NULL). */
int thx_ctx_get_reconstruction(thx_ctx* ctx, int cls, double* F, double* T, double* dir,
                               double* count);

/* ------------------------------------------------------------------------ */
/* On-disk formats (host only; no device work except thx_ctx_load_thu).      */
/* ------------------------------------------------------------------------ */

/* One particle of a .thu file: the 27 whitespace-separated fields of a line
 * (include/Database.h:12-88, read by Database, src/Database.cpp:100-300),
 * in file order. */
typedef struct thx_thu_record {
    float ctf[7];          /* voltage, defocusU, defocusV, defocusTheta, Cs,
                              amplitudeContrast, phaseShift (CTFAttr order)   */
    char particle[256];    /* "index@stack.mrcs" (1-based) or a file name      */
    char micrograph[256];
    double coord[2];       /* coordinate x, y on the micrograph                */
    int group, cls;        /* group id, class id                              */
    double quat[4];        /* rotation (w, x, y, z)                           */
    double k[3];           /* rotation concentrations k1, k2, k3              */
    double trans[2];       /* translation (pixels)                            */
    double stdTrans[2];    /* translation deviations s0, s1                   */
    double df, stdDf;      /* defocus factor and its deviation                */
    double score;
} thx_thu_record;

/* Reads the particles of a .thu file: *n receives their number; the first
 * min(n, cap) go to rec (cap 0, rec NULL: count only).  A line with fewer
 * than 27 fields is an error naming the line. */
int thx_thu_read(const char* path, thx_thu_record* rec, int cap, int* n);
/* Writes n particles (append != 0: at the end of the file) with the field
 * order and %18.9lf format of Optimiser::saveDatabase
 * (src/Optimiser.cpp:8167-8268). */
int thx_thu_write(const char* path, const thx_thu_record* rec, int n, int append);

/* MRC files (ImageFile, src/Image/ImageFile.cpp:174-300): the dimensions and
 * mode (0 int8, 1 int16, 2 float32; others THX_ERR_RANGE) of the header. */
int thx_mrc_info(const char* path, int* nx, int* ny, int* nz, int* mode);
/* Slices first .. first + count - 1 as float, nx fastest, after the 1024-byte
 * header and nsymbt bytes of extended header: out[count * ny * nx]. */
int thx_mrc_read(const char* path, int first, int count, float* out);
/* An nx x ny x nz float32 (mode 2) file with the header fillMRCHeader writes
 * (mapc/mapr/maps 1 2 3, 90 degree cell, "MAP "), the cell in Angstrom from
 * pixelSize, the density range and mean, and the little-endian stamp. */
int thx_mrc_write(const char* path, const float* data, int nx, int ny, int nz, float pixelSize);

/* Particles first .. first + count - 1 of a .thu file into the context: each
 * particle's image read from its stack under parPrefix ("index@stack" with a
 * 1-based index, as Optimiser::initImg reads them, src/Optimiser.cpp:
 * 4620-4660), then thx_ctx_set_particle_images (bgRadius, softMask; msOut as
 * there) and thx_ctx_set_ctf with the particles' CTF attributes and
 * pixelSize.  The pixel set must be set; sigRcp is left to
 * thx_ctx_init_sigma(apply). */
int thx_ctx_load_thu(thx_ctx* ctx, const char* thuPath, const char* parPrefix, int first,
                     int count, float pixelSize, float bgRadius, int softMask, float* msOut);

#ifdef __cplusplus
}
#endif
#endif
