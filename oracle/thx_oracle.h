/*
 * CPU oracle for the THUNDER expectation step (test infrastructure only).
 *
 * This library is the checker for the HIP path: only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it.  Nothing in the
 * product links it.  Every function restates one reference routine from
 * THUNDER 1.4.11 (ruanhuabin/THUNDER-Refactor-Expectation); the citation is
 * on each function.  Layout conventions follow the reference:
 *   - complex values are interleaved float pairs {re, im} (include/Complex.h);
 *   - Fourier volumes are half-complex, (V/2+1) x V x V, x fastest
 *     (include/Image/Volume.h:566-574);
 *   - rotation matrices are row-major 3x3 doubles, M[a*3+b] = M(a,b).
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed in tests/test_oracle.py; every function here has at least one
 * closed-form, invariant or brute-force check there.
 */
#ifndef THX_ORACLE_H
#define THX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* a1: pixel set within [rL, rU) of the half-plane FT of an N x N image. */
int thxo_pixel_set(int N, double rU, double rL, int pf,
                   int* iCol, int* iRow, int* iSig, int* iPxl, int cap);

/* a2: CTF values at the pixel set. */
void thxo_ctf(float* dst, float pixelSize, float voltage, float defocusU,
              float defocusV, float theta, float Cs, float amplitudeContrast,
              float phaseShift, int nCol, int nRow, const int* iCol,
              const int* iRow, int K);

/* a5: translation phase factors, dst[2k], dst[2k+1] = cos, sin. */
void thxo_translate(float* dst, float tx, float ty, int nCol, int nRow,
                    const int* iCol, const int* iRow, int K);

/* a4: rotation matrix from a unit quaternion (w, x, y, z). */
void thxo_rotmat(double* M, const double* q);

/* a6: Fourier-slice extraction by trilinear interpolation. */
void thxo_project(float* dst, const float* vol, int V, const double* M,
                  const int* iCol, const int* iRow, int K, int pf);

/* a6 in FP64 (used only by tests as a brute-force trilinear check). */
void thxo_project_f64(double* dst, const float* vol, int V, const double* M,
                      const int* iCol, const int* iRow, int K, int pf);

/* a7: pixel-major log-likelihood of n images against one prior slice. */
void thxo_logdatavs_pm(const float* dat, const float* pri, const float* ctf,
                       const float* sigRcp, int n, int m, float* result);

/* a7 as SIMD256Float computes it (AVX, 8 images per vector): equal to
 * thxo_logdatavs_pm bit for bit; -1 without AVX2. */
int thxo_logdatavs_pm_simd256(const float* dat, const float* pri, const float* ctf,
                              const float* sigRcp, int n, int m, float* result);

/* a7 in FP64 for one image-major particle; pri = S_t (.) P_r already formed. */
double thxo_logdatavs_f64(const float* dat, const float* pri, const float* ctf,
                          const float* sigRcp, int K);

/* a8: full global scan of one class, FP32, the reference loop order. */
int thxo_global_scan_f32(const float* vol, int V, int pf,
                         const int* iCol, const int* iRow, int K, int N,
                         const double* quat, int R, const double* trans, int T,
                         const float* datP, const float* ctfP,
                         const float* sigRcpP, int P,
                         const double* pR, const double* pT,
                         float* wC, float* wR, float* wT, float* baseL,
                         int nThreads);

/* a8 with the SIMD256Float likelihood sum (the reference's ENABLE_SIMD_256
 * build); -2 without AVX2. */
int thxo_global_scan_f32_simd256(const float* vol, int V, int pf,
                                 const int* iCol, const int* iRow, int K, int N,
                                 const double* quat, int R, const double* trans, int T,
                                 const float* datP, const float* ctfP,
                                 const float* sigRcpP, int P,
                                 const double* pR, const double* pT,
                                 float* wC, float* wR, float* wT, float* baseL,
                                 int nThreads);

/* a8 in FP64 with a two-pass max, scores also returned when dvp != NULL. */
int thxo_global_scan_f64(const float* vol, int V, int pf,
                         const int* iCol, const int* iRow, int K, int N,
                         const double* quat, int R, const double* trans, int T,
                         const float* datP, const float* ctfP,
                         const float* sigRcpP, int P,
                         const double* pR, const double* pT,
                         double* wC, double* wR, double* wT, double* baseL,
                         double* dvp, int nThreads);

/* a12: one local-search phase for a ragged batch, FP64 (see the .c file). */
int thxo_local_scan_f64(const float* vol, int V, int pf, const int* iCol, const int* iRow,
                        int K, int N, const double* quat, const int* rOff, const int* nR,
                        const double* trans, const int* tOff, const int* nT,
                        const double* pR, const double* pT, const float* datP,
                        const float* ctfP, const float* sigRcpP, int P, double* wC,
                        double* wR, double* wT, double* baseL, int* bestR, int* bestT,
                        int nThreads);

/* a10: systematic resampling with an injected offset u0 in [0, 1/nOut). */
int thxo_resample(const double* w, const double* u, int nIn, int nOut,
                  double u0, int* idx, double* wNew);

/* Analytic Fourier transform of an isotropic Gaussian blob sum, sampled on
 * the 2x-padded half-complex grid: the synthetic-volume closed form. */
void thxo_blob_ft(float* dst, int V, int N, const double* centres,
                  const double* amps, int nBlob, double sigma,
                  const int* ix, const int* iy, const int* iz, int n);

/* Particle update of a local phase (a9 / a11). */
int thxo_infer_acg(const double* q, int n, const double* pre, double tol, int maxIter,
                   double* Aout);
void thxo_sym4_top(const double* A, double* vec, double* evals);
int thxo_cal_vari_rot(const double* q, int n, double tol, int maxIter, double* mean, double* k);
int thxo_balance_rot(const double* q, int n, double tol, int maxIter, double* w);
void thxo_trans_stats(const double* t, int n, double* m, double* s, double* w);
double thxo_keep_peak(double* u, int n, double factor, int setFactor, int* top);

/* a9: post-scan update of one particle (class, rotations, shifts, calVari,
 * scanning minima) with the random draws given. */
int thxo_post_scan(int nCls, const double* wC, const double* base, const double* wR,
                   const double* wT, const double* pR, const double* pT, const double* gridQ,
                   const double* gridT, int R, int T, int mLR, int mLT, double u0C,
                   double pickC, double u0R, double u0T, double minStdR, double minStdT,
                   double pfS, double tol, int maxIter, int* cls, int* idxR, double* q,
                   double* wRn, int* idxT, double* t, double* wTn, double* k, double* s,
                   double* peakR);

/* F2: reconstruction insert of one image (trilinear, FP64 accumulators). */
void thxo_insert(double* F, double* T, int V, int pf, const float* src, const float* ctf,
                 const double* M, const int* iCol, const int* iRow, int K, double w);

/* a13: phase at which a particle's local search stops (-1: not within n). */
int thxo_converge(const double* k, const double* s, const double* sD, int n, int phase0,
                  int minPhase);

#ifdef __cplusplus
}
#endif
#endif

