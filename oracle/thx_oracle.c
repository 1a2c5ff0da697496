/*
 * CPU oracle for the THUNDER expectation step -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark may
 * load this library, and only as the checker.  It restates the reference
 * algorithms (THUNDER 1.4.11) in plain C; each function cites the reference
 * file:line it follows.  Arithmetic types follow the default reference build
 * (SINGLE_PRECISION=ON, so RFLOAT = float; rotations and particle weights are
 * double), except the *_f64 variants, which are the FP64 references.
 *
 * Parity pinning: the reference ships no golden vectors or tests for this
 * path, so each function is pinned by closed forms, invariants and brute
 * force in tests/test_oracle.py (named there per function).
 */
#include "thx_oracle.h"

#include <float.h>
#include <immintrin.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* AROUND(a) = (int)rint(a), NORM = gsl_hypot (include/Functions/Functions.h:30-70). */
static int around(double a) { return (int)rint(a); }

/*
 * Optimiser::allocPreCalIdx, src/Optimiser.cpp:7942-7992.
 * Loop IMAGE_FOR_PIXEL_R_FT(rU + 1) (include/Image/Image.h:68-70): j outer in
 * [-r, r), i inner in [0, r]; skip (i == 0, j < 0); keep rL^2 <= i^2+j^2 < rU^2
 * and rL <= rint(hypot) < rU.  iPxl = iFTHalf(i, j) of an N x N image
 * (include/Image/Image.h:426-432).  Returns K, or -1 if cap is exceeded.
 */
int thxo_pixel_set(int N, double rU, double rL, int pf,
                   int* iCol, int* iRow, int* iSig, int* iPxl, int cap)
{
    (void)pf;
    float rUf = (float)rU, rLf = (float)rL;       /* RFLOAT arguments */
    float r = rUf + 1;
    float rU2 = rUf * rUf, rL2 = rLf * rLf;
    int nColFT = N / 2 + 1;
    int K = 0;
    for (int j = (int)(-r); j < r; j++)
        for (int i = 0; i <= r; i++) {
            if (i == 0 && j < 0) continue;
            float u = (float)((double)i * i + (double)j * j);
            if (u < rU2 && u >= rL2) {
                int v = around(hypot((double)i, (double)j));
                if (v < rUf && v >= rLf) {
                    if (K >= cap) return -1;
                    if (iPxl) iPxl[K] = (j >= 0 ? j : j + N) * nColFT + i;
                    iCol[K] = i;
                    iRow[K] = j;
                    if (iSig) iSig[K] = v;
                    K++;
                }
            }
        }
    return K;
}

/*
 * CTF(RFLOAT* dst, ...), src/CTF.cpp:109-145.  Units: V, Angstrom, rad.
 * lambda = 12.2643247 / sqrt(V (1 + 0.978466e-6 V)); K1 = pi lambda;
 * K2 = pi/2 Cs lambda^3; chi = K1 df u^2 + K2 u^4 - phaseShift;
 * ctf = -sqrt(1 - A^2) sin(chi) + A cos(chi).
 */
void thxo_ctf(float* dst, float pixelSize, float voltage, float defocusU,
              float defocusV, float theta, float Cs, float amplitudeContrast,
              float phaseShift, int nCol, int nRow, const int* iCol,
              const int* iRow, int K)
{
    float lambda = (float)(12.2643247 / sqrt(voltage * (1 + voltage * 0.978466e-6)));
    float w1 = sqrtf(1 - amplitudeContrast * amplitudeContrast);
    float w2 = amplitudeContrast;
    float K1 = (float)(M_PI * lambda);
    float K2 = (float)(M_PI / 2 * Cs * ((double)lambda * lambda * lambda));
    for (int i = 0; i < K; i++) {
        float u = (float)hypot(iCol[i] / (pixelSize * nCol), iRow[i] / (pixelSize * nRow));
        float angle = (float)(atan2((double)iRow[i], (double)iCol[i]) - theta);
        float defocus = -(defocusU + defocusV + (defocusU - defocusV) * cosf(2 * angle)) / 2;
        float u2 = u * u;
        float ki = K1 * defocus * u2 + K2 * (u2 * u2) - phaseShift;
        dst[i] = -w1 * sinf(ki) + w2 * cosf(ki);
    }
}

/*
 * translate(Complex* dst, tx, ty, nCol, nRow, iCol, iRow, nPxl),
 * src/Image/ImageFunctions.cpp:233-251 with COMPLEX_POLAR (include/Complex.h:33-38):
 * dst = exp(-i 2 pi (iCol tx / nCol + iRow ty / nRow)).
 */
void thxo_translate(float* dst, float tx, float ty, int nCol, int nRow,
                    const int* iCol, const int* iRow, int K)
{
    float rCol = tx / nCol, rRow = ty / nRow;
    for (int i = 0; i < K; i++) {
        float phase = (float)(2 * M_PI) * (iCol[i] * rCol + iRow[i] * rRow);
        dst[2 * i] = cosf(-phase);
        dst[2 * i + 1] = sinf(-phase);
    }
}

/*
 * rotate3D(dmat33&, const dvec4&), src/Geometry/Euler.cpp:181-189:
 * R = I + 2 q0 A + 2 A^2 with A = [q]x (cross-product matrix of (q1, q2, q3)).
 */
void thxo_rotmat(double* M, const double* q)
{
    double A[9] = {0, -q[3], q[2], q[3], 0, -q[1], -q[2], q[1], 0};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double aa = 0;
            for (int c = 0; c < 3; c++) aa += A[a * 3 + c] * A[c * 3 + b];
            M[a * 3 + b] = (a == b ? 1.0 : 0.0) + 2 * q[0] * A[a * 3 + b] + 2 * aa;
        }
}

/* Volume::iFTHalf (include/Image/Volume.h:566-574): wrap negative y, z. */
static size_t ifthalf(int V, int i, int j, int k)
{
    int nColFT = V / 2 + 1;
    return (size_t)(k >= 0 ? k : k + V) * nColFT * V + (size_t)(j >= 0 ? j : j + V) * nColFT + i;
}

/*
 * Projector::project(Complex*, const dmat33&, iCol, iRow, nPxl),
 * src/Projector.cpp:356-374 -> Volume::getByInterpolationFT,
 * src/Image/Volume.cpp:308-332 -> conjHalf (include/Image/Volume.h:139-151)
 * -> WG_TRI_INTERP_LINEAR (include/Functions/Interpolation.h:195-208)
 * -> Volume::getFTHalf(w, x0), src/Image/Volume.cpp:481-553.
 * Coordinates are formed in FP64 and passed as RFLOAT (float); x < 0 is
 * mirrored and conjugated; corners are fetched with y/z wrap; the 8 products
 * are summed in the order w[k][j][i] flattened (k slowest).
 */
void thxo_project(float* dst, const float* vol, int V, const double* M,
                  const int* iCol, const int* iRow, int K, int pf)
{
    for (int n = 0; n < K; n++) {
        double c0 = (double)(iCol[n] * pf), c1 = (double)(iRow[n] * pf);
        float x[3];
        for (int a = 0; a < 3; a++) x[a] = (float)(M[a * 3 + 0] * c0 + M[a * 3 + 1] * c1);
        int conj = 0;
        if (x[0] < 0) { x[0] = -x[0]; x[1] = -x[1]; x[2] = -x[2]; conj = 1; }
        int x0[3];
        float v[3][2];
        for (int a = 0; a < 3; a++) {
            x0[a] = (int)floorf(x[a]);
            float xd = x[a] - x0[a];
            v[a][0] = 1 - xd;
            v[a][1] = xd;
        }
        float re = 0, im = 0;
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) {
                    float w = v[0][i] * v[1][j] * v[2][k];
                    size_t idx = ifthalf(V, x0[0] + i, x0[1] + j, x0[2] + k);
                    re += vol[2 * idx] * w;
                    im += vol[2 * idx + 1] * w;
                }
        dst[2 * n] = re;
        dst[2 * n + 1] = conj ? -im : im;
    }
}

void thxo_project_f64(double* dst, const float* vol, int V, const double* M,
                      const int* iCol, const int* iRow, int K, int pf)
{
    for (int n = 0; n < K; n++) {
        double c0 = (double)(iCol[n] * pf), c1 = (double)(iRow[n] * pf);
        double x[3];
        for (int a = 0; a < 3; a++) x[a] = M[a * 3 + 0] * c0 + M[a * 3 + 1] * c1;
        int conj = 0;
        if (x[0] < 0) { x[0] = -x[0]; x[1] = -x[1]; x[2] = -x[2]; conj = 1; }
        int x0[3];
        double v[3][2];
        for (int a = 0; a < 3; a++) {
            x0[a] = (int)floor(x[a]);
            double xd = x[a] - x0[a];
            v[a][0] = 1 - xd;
            v[a][1] = xd;
        }
        double re = 0, im = 0;
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) {
                    double w = v[0][i] * v[1][j] * v[2][k];
                    size_t idx = ifthalf(V, x0[0] + i, x0[1] + j, x0[2] + k);
                    re += vol[2 * idx] * w;
                    im += vol[2 * idx + 1] * w;
                }
        dst[2 * n] = re;
        dst[2 * n + 1] = conj ? -im : im;
    }
}

/*
 * logDataVSPrior_m_n_huabin, src/Optimiser.cpp:9666-9708 (scalar form of
 * SIMD256Float :8957-9041): pixel-major arrays idx = i * n + j;
 * result[j] += sigRcp * |dat - ctf * pri[i]|^2.
 */
/* no contraction into FMA in the two forms of this sum, so that they agree bit
 * for bit */
#pragma GCC push_options
#pragma GCC optimize("fp-contract=off")
void thxo_logdatavs_pm(const float* dat, const float* pri, const float* ctf,
                       const float* sigRcp, int n, int m, float* result)
{
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            size_t idx = (size_t)i * n + j;
            float cr = ctf[idx] * pri[2 * i], ci = ctf[idx] * pri[2 * i + 1];
            float dr = dat[2 * idx] - cr, di = dat[2 * idx + 1] - ci;
            float t = dr * dr + di * di;
            result[j] += t * sigRcp[idx];
        }
}

/*
 * SIMD256Float (src/Optimiser.cpp:8957-9041), the AVX form of the same sum
 * the reference's CPU expectation runs with ENABLE_SIMD_256: per pixel i the
 * images j are taken 8 at a time, with the same operations in the same order
 * per image (ctf * pri, dat - that, squares, their sum, times sigRcp, added to
 * result[j]) and the same scalar tail, so every result[j] equals
 * thxo_logdatavs_pm's bit for bit (neither is contracted into FMA).
 * Returns 0, or -1 when the CPU has no AVX2 (nothing done).
 */
__attribute__((target("avx2"))) static void logdatavs_pm_avx256(
    const float* dat, const float* pri, const float* ctf, const float* sigRcp, int n, int m,
    float* result)
{
    for (int i = 0; i < m; i++) {
        const __m256 pr = _mm256_set1_ps(pri[2 * i]), pim = _mm256_set1_ps(pri[2 * i + 1]);
        int j = 0;
        for (; j <= n - 8; j += 8) {
            const size_t idx = (size_t)i * n + j;
            const __m256 c = _mm256_loadu_ps(ctf + idx);
            /* (re, im) pairs of 8 images -> 8 reals and 8 imaginaries */
            const __m256 a = _mm256_loadu_ps(dat + 2 * idx);
            const __m256 b = _mm256_loadu_ps(dat + 2 * idx + 8);
            __m256 re = _mm256_shuffle_ps(a, b, _MM_SHUFFLE(2, 0, 2, 0));
            __m256 im = _mm256_shuffle_ps(a, b, _MM_SHUFFLE(3, 1, 3, 1));
            re = _mm256_castpd_ps(_mm256_permute4x64_pd(_mm256_castps_pd(re),
                                                        _MM_SHUFFLE(3, 1, 2, 0)));
            im = _mm256_castpd_ps(_mm256_permute4x64_pd(_mm256_castps_pd(im),
                                                        _MM_SHUFFLE(3, 1, 2, 0)));
            __m256 tr = _mm256_mul_ps(c, pr), ti = _mm256_mul_ps(c, pim);
            tr = _mm256_sub_ps(re, tr);
            ti = _mm256_sub_ps(im, ti);
            tr = _mm256_mul_ps(tr, tr);
            ti = _mm256_mul_ps(ti, ti);
            tr = _mm256_add_ps(tr, ti);
            tr = _mm256_mul_ps(tr, _mm256_loadu_ps(sigRcp + idx));
            _mm256_storeu_ps(result + j, _mm256_add_ps(_mm256_loadu_ps(result + j), tr));
        }
        for (; j < n; j++) {
            const size_t idx = (size_t)i * n + j;
            float cr = ctf[idx] * pri[2 * i], ci = ctf[idx] * pri[2 * i + 1];
            float dr = dat[2 * idx] - cr, di = dat[2 * idx + 1] - ci;
            float t = dr * dr + di * di;
            result[j] += t * sigRcp[idx];
        }
    }
}

int thxo_logdatavs_pm_simd256(const float* dat, const float* pri, const float* ctf,
                              const float* sigRcp, int n, int m, float* result)
{
    if (!__builtin_cpu_supports("avx2")) return -1;
    logdatavs_pm_avx256(dat, pri, ctf, sigRcp, n, m, result);
    return 0;
}
#pragma GCC pop_options

/* FP64 reference of the same sum for one image-major particle. */
double thxo_logdatavs_f64(const float* dat, const float* pri, const float* ctf,
                          const float* sigRcp, int K)
{
    double s = 0;
    for (int i = 0; i < K; i++) {
        double cr = (double)ctf[i] * pri[2 * i], ci = (double)ctf[i] * pri[2 * i + 1];
        double dr = dat[2 * i] - cr, di = dat[2 * i + 1] - ci;
        s += (dr * dr + di * di) * sigRcp[i];
    }
    return s;
}

/*
 * The hot loop of Optimiser::doGlobalSearch, src/Optimiser.cpp:711-915, for
 * one class (t = 0, k = 1): traP per translation (:711-725); OpenMP over
 * rotations (:759); per rotation project (:782), per translation
 * pri = tra (.) rot (:793-794), dvp over all particles pixel-major (:801-827),
 * then per particle under its lock the streaming baseline and
 * wC += w pR pT, wR += w pT, wT += w pR (:835-895).  Inputs are image-major
 * (l*K + i) as the GPU interface takes them; they are transposed to the
 * pixel-major order the CPU path uses (allocPreCal, :7994-8121).
 * wC[P], wR[P*R], wT[P*T], baseL[P].
 */
static int global_scan_f32(const float* vol, int V, int pf,
                           const int* iCol, const int* iRow, int K, int N,
                           const double* quat, int R, const double* trans, int T,
                           const float* datP, const float* ctfP,
                           const float* sigRcpP, int P,
                           const double* pR, const double* pT,
                           float* wC, float* wR, float* wT, float* baseL,
                           int nThreads, int simd)
{
    if (simd && !__builtin_cpu_supports("avx2")) return -2;
#ifdef _OPENMP
    if (nThreads > 0) omp_set_num_threads(nThreads);
#else
    (void)nThreads;
#endif
    float* datPM = (float*)malloc(sizeof(float) * 2 * (size_t)P * K);
    float* ctfPM = (float*)malloc(sizeof(float) * (size_t)P * K);
    float* sigPM = (float*)malloc(sizeof(float) * (size_t)P * K);
    float* traP = (float*)malloc(sizeof(float) * 2 * (size_t)T * K);
    if (!datPM || !ctfPM || !sigPM || !traP) {
        free(datPM); free(ctfPM); free(sigPM); free(traP);
        return -1;
    }
    for (int l = 0; l < P; l++)
        for (int i = 0; i < K; i++) {
            size_t s = (size_t)l * K + i, d = (size_t)i * P + l;
            datPM[2 * d] = datP[2 * s];
            datPM[2 * d + 1] = datP[2 * s + 1];
            ctfPM[d] = ctfP[s];
            sigPM[d] = sigRcpP[s];
        }
    for (int m = 0; m < T; m++)
        thxo_translate(traP + (size_t)2 * m * K, (float)trans[2 * m], (float)trans[2 * m + 1],
                       N, N, iCol, iRow, K);

    memset(wC, 0, sizeof(float) * P);
    memset(wR, 0, sizeof(float) * (size_t)P * R);
    memset(wT, 0, sizeof(float) * (size_t)P * T);
#ifdef _OPENMP
    omp_lock_t* mtx = (omp_lock_t*)malloc(sizeof(omp_lock_t) * P);
    for (int l = 0; l < P; l++) omp_init_lock(&mtx[l]);
#endif
    for (int l = 0; l < P; l++) baseL[l] = NAN;

    #pragma omp parallel
    {
        float* priRotP = (float*)malloc(sizeof(float) * 2 * K);
        float* priAllP = (float*)malloc(sizeof(float) * 2 * K);
        float* res = (float*)malloc(sizeof(float) * P);
        #pragma omp for schedule(dynamic)
        for (int m = 0; m < R; m++) {
            double M[9];
            thxo_rotmat(M, quat + 4 * m);
            thxo_project(priRotP, vol, V, M, iCol, iRow, K, pf);
            for (int n = 0; n < T; n++) {
                const float* tra = traP + (size_t)2 * n * K;
                for (int i = 0; i < K; i++) {
                    float ar = tra[2 * i], ai = tra[2 * i + 1];
                    float br = priRotP[2 * i], bi = priRotP[2 * i + 1];
                    priAllP[2 * i] = ar * br - ai * bi;
                    priAllP[2 * i + 1] = ar * bi + ai * br;
                }
                memset(res, 0, sizeof(float) * P);
                if (simd)
                    logdatavs_pm_avx256(datPM, priAllP, ctfPM, sigPM, P, K, res);
                else
                    thxo_logdatavs_pm(datPM, priAllP, ctfPM, sigPM, P, K, res);
                for (int l = 0; l < P; l++) {
#ifdef _OPENMP
                    omp_set_lock(&mtx[l]);
#endif
                    if (isnan(baseL[l]))
                        baseL[l] = res[l];
                    else if (res[l] > baseL[l]) {
                        float offset = res[l] - baseL[l];
                        float nf = expf(-offset);
                        wC[l] *= nf;
                        for (int r = 0; r < R; r++) wR[(size_t)l * R + r] *= nf;
                        for (int t = 0; t < T; t++) wT[(size_t)l * T + t] *= nf;
                        baseL[l] += offset;
                    }
                    float w = expf(res[l] - baseL[l]);
                    wC[l] += (float)(w * (pR[m] * pT[n]));
                    wR[(size_t)l * R + m] += (float)(w * pT[n]);
                    wT[(size_t)l * T + n] += (float)(w * pR[m]);
#ifdef _OPENMP
                    omp_unset_lock(&mtx[l]);
#endif
                }
            }
        }
        free(priRotP); free(priAllP); free(res);
    }
#ifdef _OPENMP
    for (int l = 0; l < P; l++) omp_destroy_lock(&mtx[l]);
    free(mtx);
#endif
    free(datPM); free(ctfPM); free(sigPM); free(traP);
    return 0;
}

int thxo_global_scan_f32(const float* vol, int V, int pf,
                         const int* iCol, const int* iRow, int K, int N,
                         const double* quat, int R, const double* trans, int T,
                         const float* datP, const float* ctfP,
                         const float* sigRcpP, int P,
                         const double* pR, const double* pT,
                         float* wC, float* wR, float* wT, float* baseL,
                         int nThreads)
{
    return global_scan_f32(vol, V, pf, iCol, iRow, K, N, quat, R, trans, T, datP, ctfP, sigRcpP,
                           P, pR, pT, wC, wR, wT, baseL, nThreads, 0);
}

/* The same scan with the likelihood sum of SIMD256Float (the reference's
 * ENABLE_SIMD_256 build, src/Optimiser.cpp:811-818): -2 when the CPU has no
 * AVX2. */
int thxo_global_scan_f32_simd256(const float* vol, int V, int pf,
                                 const int* iCol, const int* iRow, int K, int N,
                                 const double* quat, int R, const double* trans, int T,
                                 const float* datP, const float* ctfP,
                                 const float* sigRcpP, int P,
                                 const double* pR, const double* pT,
                                 float* wC, float* wR, float* wT, float* baseL,
                                 int nThreads)
{
    return global_scan_f32(vol, V, pf, iCol, iRow, K, N, quat, R, trans, T, datP, ctfP, sigRcpP,
                           P, pR, pT, wC, wR, wT, baseL, nThreads, 1);
}

/*
 * FP64 reference of the same marginals: dvp[l][m][n] in FP64 from the FP64
 * projection and phases, baseline = max over (m, n) (two passes, so the
 * result does not depend on the streaming order of :835-895), then
 * wC = sum w pR pT, wR[m] = sum_n w pT, wT[n] = sum_m w pR.
 * dvp (optional) is P x R x T.
 */
int thxo_global_scan_f64(const float* vol, int V, int pf,
                         const int* iCol, const int* iRow, int K, int N,
                         const double* quat, int R, const double* trans, int T,
                         const float* datP, const float* ctfP,
                         const float* sigRcpP, int P,
                         const double* pR, const double* pT,
                         double* wC, double* wR, double* wT, double* baseL,
                         double* dvp, int nThreads)
{
#ifdef _OPENMP
    if (nThreads > 0) omp_set_num_threads(nThreads);
#else
    (void)nThreads;
#endif
    double* rot = (double*)malloc(sizeof(double) * 2 * (size_t)R * K);
    double* tra = (double*)malloc(sizeof(double) * 2 * (size_t)T * K);
    double* sc = (double*)malloc(sizeof(double) * (size_t)P * R * T);
    if (!rot || !tra || !sc) { free(rot); free(tra); free(sc); return -1; }
    #pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < R; m++) {
        double M[9];
        thxo_rotmat(M, quat + 4 * m);
        thxo_project_f64(rot + (size_t)2 * m * K, vol, V, M, iCol, iRow, K, pf);
    }
    for (int n = 0; n < T; n++)
        for (int i = 0; i < K; i++) {
            double ph = 2 * M_PI * (iCol[i] * trans[2 * n] / N + iRow[i] * trans[2 * n + 1] / N);
            tra[(size_t)2 * (n * K + i)] = cos(-ph);
            tra[(size_t)2 * (n * K + i) + 1] = sin(-ph);
        }
    #pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < P; l++) {
        const float* d = datP + (size_t)2 * l * K;
        const float* c = ctfP + (size_t)l * K;
        const float* s = sigRcpP + (size_t)l * K;
        for (int m = 0; m < R; m++)
            for (int n = 0; n < T; n++) {
                const double* pr = rot + (size_t)2 * m * K;
                const double* tr = tra + (size_t)2 * n * K;
                double acc = 0;
                for (int i = 0; i < K; i++) {
                    double qr = tr[2 * i] * pr[2 * i] - tr[2 * i + 1] * pr[2 * i + 1];
                    double qi = tr[2 * i] * pr[2 * i + 1] + tr[2 * i + 1] * pr[2 * i];
                    double er = d[2 * i] - c[i] * qr, ei = d[2 * i + 1] - c[i] * qi;
                    acc += (er * er + ei * ei) * s[i];
                }
                sc[((size_t)l * R + m) * T + n] = acc;
            }
        double mx = -INFINITY;
        for (size_t q = 0; q < (size_t)R * T; q++)
            if (sc[(size_t)l * R * T + q] > mx) mx = sc[(size_t)l * R * T + q];
        baseL[l] = mx;
        double c0 = 0;
        for (int m = 0; m < R; m++) wR[(size_t)l * R + m] = 0;
        for (int n = 0; n < T; n++) wT[(size_t)l * T + n] = 0;
        for (int m = 0; m < R; m++)
            for (int n = 0; n < T; n++) {
                double w = exp(sc[((size_t)l * R + m) * T + n] - mx);
                c0 += w * pR[m] * pT[n];
                wR[(size_t)l * R + m] += w * pT[n];
                wT[(size_t)l * T + n] += w * pR[m];
            }
        wC[l] = c0;
    }
    if (dvp) memcpy(dvp, sc, sizeof(double) * (size_t)P * R * T);
    free(rot); free(tra); free(sc);
    return 0;
}

/*
 * One local-search phase for a batch of particles, FP64 reference
 * (Optimiser::updateSP, src/Optimiser.cpp:1139-1437, one class, no CTF
 * search): particle l scores its own rotations quat[rOff[l] .. + nR[l]) and
 * translations trans[tOff[l] .. + nT[l]) with priors pR, pT (same offsets):
 * base = max dvp, wC = sum w pR pT, wR[r] = sum_t w pT, wT[t] = sum_r w pR
 * (the reference's streaming baseline of :1330-1349 in two-pass form, with
 * wC(iC) = 1).  wR and wT use the rOff / tOff offsets; wC, base per particle.
 */
int thxo_local_scan_f64(const float* vol, int V, int pf, const int* iCol, const int* iRow,
                        int K, int N, const double* quat, const int* rOff, const int* nR,
                        const double* trans, const int* tOff, const int* nT,
                        const double* pR, const double* pT, const float* datP,
                        const float* ctfP, const float* sigRcpP, int P, double* wC,
                        double* wR, double* wT, double* baseL, int* bestR, int* bestT,
                        int nThreads)
{
#ifdef _OPENMP
    if (nThreads > 0) omp_set_num_threads(nThreads);
#else
    (void)nThreads;
#endif
    int fail = 0;
    #pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < P; l++) {
        const int R = nR[l], T = nT[l];
        double* rot = (double*)malloc(sizeof(double) * 2 * (size_t)R * K);
        double* tra = (double*)malloc(sizeof(double) * 2 * (size_t)T * K);
        double* sc = (double*)malloc(sizeof(double) * (size_t)R * T);
        if (!rot || !tra || !sc) {
            free(rot); free(tra); free(sc);
            #pragma omp atomic write
            fail = 1;
            continue;
        }
        for (int m = 0; m < R; m++) {
            double M[9];
            thxo_rotmat(M, quat + 4 * (size_t)(rOff[l] + m));
            thxo_project_f64(rot + (size_t)2 * m * K, vol, V, M, iCol, iRow, K, pf);
        }
        for (int n = 0; n < T; n++) {
            const double* t = trans + 2 * (size_t)(tOff[l] + n);
            for (int i = 0; i < K; i++) {
                double ph = 2 * M_PI * (iCol[i] * t[0] / N + iRow[i] * t[1] / N);
                tra[(size_t)2 * (n * K + i)] = cos(-ph);
                tra[(size_t)2 * (n * K + i) + 1] = sin(-ph);
            }
        }
        const float* d = datP + (size_t)2 * l * K;
        const float* c = ctfP + (size_t)l * K;
        const float* s = sigRcpP + (size_t)l * K;
        double mx = -INFINITY;
        int br = -1, bt = -1;
        for (int m = 0; m < R; m++)
            for (int n = 0; n < T; n++) {
                const double* pr = rot + (size_t)2 * m * K;
                const double* tr = tra + (size_t)2 * n * K;
                double acc = 0;
                for (int i = 0; i < K; i++) {
                    double qr = tr[2 * i] * pr[2 * i] - tr[2 * i + 1] * pr[2 * i + 1];
                    double qi = tr[2 * i] * pr[2 * i + 1] + tr[2 * i + 1] * pr[2 * i];
                    double er = d[2 * i] - c[i] * qr, ei = d[2 * i + 1] - c[i] * qi;
                    acc += (er * er + ei * ei) * s[i];
                }
                sc[(size_t)m * T + n] = acc;
                if (acc > mx) { mx = acc; br = m; bt = n; }
            }
        baseL[l] = mx;
        if (bestR) bestR[l] = br;
        if (bestT) bestT[l] = bt;
        double c0 = 0;
        double* wr = wR + rOff[l];
        double* wt = wT + tOff[l];
        for (int m = 0; m < R; m++) wr[m] = 0;
        for (int n = 0; n < T; n++) wt[n] = 0;
        for (int m = 0; m < R; m++)
            for (int n = 0; n < T; n++) {
                double w = exp(sc[(size_t)m * T + n] - mx);
                double a = pR[rOff[l] + m], b = pT[tOff[l] + n];
                c0 += w * a * b;
                wr[m] += w * b;
                wt[n] += w * a;
            }
        wC[l] = c0;
        free(rot); free(tra); free(sc);
    }
    return fail ? -1 : 0;
}

/*
 * Particle::resample(n, pt), src/Particle.cpp:1275-1462, one sample type,
 * without the random shuffle (:2173-2271) and with u0 injected:
 * w <- w (.) u, normalise, cdf (d_cumsum) / cdf[last], u_j = u0 + j / nOut,
 * advance i while u_j > cdf[i]; new weight 1 / u_i (PARTICLE_PRIOR_ONE,
 * include/Config.h:63), then normW (src/Particle.cpp:813-819) normalises.
 * Returns 0, or -1 on bad input.
 */
int thxo_resample(const double* w, const double* u, int nIn, int nOut,
                  double u0, int* idx, double* wNew)
{
    if (nIn <= 0 || nOut <= 0) return -1;
    double* cdf = (double*)malloc(sizeof(double) * nIn);
    if (!cdf) return -1;
    double sum = 0;
    for (int i = 0; i < nIn; i++) sum += w[i] * u[i];
    double acc = 0;
    for (int i = 0; i < nIn; i++) {
        acc += (w[i] * u[i]) / sum;
        cdf[i] = acc;
    }
    double last = cdf[nIn - 1];
    for (int i = 0; i < nIn; i++) cdf[i] /= last;
    int i = 0;
    double wsum = 0;
    for (int j = 0; j < nOut; j++) {
        double uj = u0 + j * 1.0 / nOut;
        while (uj > cdf[i] && i < nIn - 1) i++;
        idx[j] = i;
        wNew[j] = 1.0 / u[i];
        wsum += wNew[j];
    }
    for (int j = 0; j < nOut; j++) wNew[j] /= wsum;
    free(cdf);
    return 0;
}

/*
 * Closed-form FT of f(x) = sum_b a_b exp(-|x - c_b|^2 / (2 s^2)) (x in voxels
 * of the N^3 box, origin at the box centre) sampled at padded-grid index m of
 * a V = pf N transform:
 * F(m) = sum_b a_b (2 pi)^{3/2} s^3 exp(-2 pi^2 s^2 |m|^2 / V^2) exp(-2 pi i m.c_b / V).
 */
void thxo_blob_ft(float* dst, int V, int N, const double* centres,
                  const double* amps, int nBlob, double sigma,
                  const int* ix, const int* iy, const int* iz, int n)
{
    (void)N;
    double g0 = pow(2 * M_PI, 1.5) * sigma * sigma * sigma;
    for (int q = 0; q < n; q++) {
        double m2 = (double)ix[q] * ix[q] + (double)iy[q] * iy[q] + (double)iz[q] * iz[q];
        double g = g0 * exp(-2 * M_PI * M_PI * sigma * sigma * m2 / ((double)V * V));
        double re = 0, im = 0;
        for (int b = 0; b < nBlob; b++) {
            double ph = -2 * M_PI * (ix[q] * centres[3 * b] + iy[q] * centres[3 * b + 1] +
                                     iz[q] * centres[3 * b + 2]) / V;
            re += amps[b] * cos(ph);
            im += amps[b] * sin(ph);
        }
        dst[2 * q] = (float)(g * re);
        dst[2 * q + 1] = (float)(g * im);
    }
}

/* ------------------------------------------------------------------------
 * Particle update of one local phase (a9 / a11 / a13).
 * ---------------------------------------------------------------------- */

/* 4x4 inverse and determinant by Gauss-Jordan with partial pivoting. */
static int inv4(const double* A, double* inv, double* det)
{
    double m[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++)
            m[i][j] = j < 4 ? A[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
    double d = 1.0;
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(m[r][c]) > fabs(m[p][c])) p = r;
        if (m[p][c] == 0.0) return -1;
        if (p != c) {
            for (int j = 0; j < 8; j++) { double t = m[p][j]; m[p][j] = m[c][j]; m[c][j] = t; }
            d = -d;
        }
        d *= m[c][c];
        double pv = m[c][c];
        for (int j = 0; j < 8; j++) m[c][j] /= pv;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                double f = m[r][c];
                for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
            }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) inv[4 * i + j] = m[i][j + 4];
    if (det) *det = d;
    return 0;
}

static void qmul_o(const double* a, const double* b, double* o)
{
    double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w; o[1] = x; o[2] = y; o[3] = z;
}

/*
 * inferACG(dmat44&, const dmat4&), src/Geometry/DirectionalStat.cpp:87-139:
 * A <- I; repeat B = (4 / sum 1/u_i) sum q_i q_i^T / u_i with u_i = q_i^T A^-1 q_i,
 * until sum |A - B| <= tol (1e-3), returning the last A (the one before B).
 * pre (optional): every sample is first left-multiplied by it (the frame
 * change of Particle::calVari, src/Particle.cpp:1043-1052).  Returns the
 * iteration count, or -1 (singular) / -2 (maxIter reached).
 */
int thxo_infer_acg(const double* q, int n, const double* pre, double tol, int maxIter,
                   double* Aout)
{
    double A[16], B[16], Ai[16];
    for (int i = 0; i < 16; i++) B[i] = (i % 5 == 0) ? 1.0 : 0.0;
    int it = 0;
    for (;;) {
        memcpy(A, B, sizeof(A));
        if (inv4(A, Ai, NULL)) return -1;
        memset(B, 0, sizeof(B));
        double nf = 0;
        for (int i = 0; i < n; i++) {
            double x[4];
            if (pre) qmul_o(pre, q + 4 * i, x);
            else memcpy(x, q + 4 * i, sizeof(x));
            double v[4] = {0, 0, 0, 0};
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) v[a] += Ai[4 * a + b] * x[b];
            double u = x[0] * v[0] + x[1] * v[1] + x[2] * v[2] + x[3] * v[3];
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) B[4 * a + b] += x[a] * x[b] / u;
            nf += 1.0 / u;
        }
        for (int i = 0; i < 16; i++) B[i] *= 4.0 / nf;
        it++;
        double diff = 0;
        for (int i = 0; i < 16; i++) diff += fabs(A[i] - B[i]);
        if (diff <= tol) break;
        if (it >= maxIter) { memcpy(Aout, A, sizeof(A)); return -2; }
    }
    memcpy(Aout, A, sizeof(A));
    return it;
}

/* Eigenvector of the largest eigenvalue of a symmetric 4x4 (cyclic Jacobi),
 * normalised, sign fixed so that its largest-magnitude entry is positive
 * (inferACG(dvec4&, ...), src/Geometry/DirectionalStat.cpp:218-244; the sign
 * of an eigenvector is arbitrary and q, -q are the same rotation). */
void thxo_sym4_top(const double* Ain, double* vec, double* evals)
{
    double A[4][4], V[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) { A[i][j] = Ain[4 * i + j]; V[i][j] = i == j; }
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) off += A[a][b] * A[a][b];
        if (off < 1e-40) break;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) {
                if (fabs(A[a][b]) < 1e-300) continue;
                double th = (A[b][b] - A[a][a]) / (2 * A[a][b]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; k++) {
                    double x = A[k][a], y = A[k][b];
                    A[k][a] = c * x - s * y; A[k][b] = s * x + c * y;
                }
                for (int k = 0; k < 4; k++) {
                    double x = A[a][k], y = A[b][k];
                    A[a][k] = c * x - s * y; A[b][k] = s * x + c * y;
                }
                for (int k = 0; k < 4; k++) {
                    double x = V[k][a], y = V[k][b];
                    V[k][a] = c * x - s * y; V[k][b] = s * x + c * y;
                }
            }
    }
    int top = 0;
    for (int a = 1; a < 4; a++) if (A[a][a] > A[top][top]) top = a;
    int big = 0;
    for (int a = 1; a < 4; a++) if (fabs(V[a][top]) > fabs(V[big][top])) big = a;
    double sg = V[big][top] < 0 ? -1.0 : 1.0, nrm = 0;
    for (int a = 0; a < 4; a++) nrm += V[a][top] * V[a][top];
    nrm = sqrt(nrm);
    for (int a = 0; a < 4; a++) vec[a] = sg * V[a][top] / nrm;
    if (evals) for (int a = 0; a < 4; a++) evals[a] = A[a][a];
}

/*
 * Particle::calVari(PAR_R), 3-D branch with PARTICLE_ROT_MEAN_USING_STAT_CAL_VARI
 * (src/Particle.cpp:1002-1081, include/Config.h:77): mean = inferACG mean of the
 * samples; k_a = A'(a,a) / A'(0,0) of inferACG of conj(mean) q
 * (inferACG(k1,k2,k3,...), src/Geometry/DirectionalStat.cpp:178-216).
 */
int thxo_cal_vari_rot(const double* q, int n, double tol, int maxIter, double* mean, double* k)
{
    double A[16];
    int it = thxo_infer_acg(q, n, NULL, tol, maxIter, A);
    if (it < 0) return it;
    thxo_sym4_top(A, mean, NULL);
    double mc[4] = {mean[0], -mean[1], -mean[2], -mean[3]};
    it = thxo_infer_acg(q, n, mc, tol, maxIter, A);
    if (it < 0) return it;
    for (int a = 0; a < 3; a++) k[a] = A[5 * (a + 1)] / A[0];
    return 0;
}

/*
 * Particle::balanceWeight(PAR_R), 3-D (src/Particle.cpp:2301-2311) with normW
 * (:813-819): w_i = 1 / pdfACG(q_i, A), pdfACG(x, A) = det(A)^-1/2 (x^T A^-1 x)^-2
 * (src/Geometry/DirectionalStat.cpp:13-18), A = inferACG of the samples; then
 * normalised to sum 1.
 */
int thxo_balance_rot(const double* q, int n, double tol, int maxIter, double* w)
{
    double A[16], Ai[16], det;
    int it = thxo_infer_acg(q, n, NULL, tol, maxIter, A);
    if (it < 0) return it;
    if (inv4(A, Ai, &det)) return -1;
    double s = 0;
    for (int i = 0; i < n; i++) {
        const double* x = q + 4 * i;
        double u = 0;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) u += x[a] * Ai[4 * a + b] * x[b];
        w[i] = sqrt(det) * u * u;
        s += w[i];
    }
    for (int i = 0; i < n; i++) w[i] /= s;
    return 0;
}

/*
 * Translation statistics of one particle's samples t[n][2]: means m, sample
 * standard deviations s (n - 1 denominators, gsl_stats_sd / gsl_stats_sd_m) as
 * Particle::calVari(PAR_T) (src/Particle.cpp:1083-1097, zero-mean switch off)
 * and balanceWeight(PAR_T) (:2313-2344, rho = 0): w_i = 1 / N2(t_i - m; s0, s1)
 * normalised.  w may be NULL.
 */
void thxo_trans_stats(const double* t, int n, double* m, double* s, double* w)
{
    double m0 = 0, m1 = 0;
    for (int i = 0; i < n; i++) { m0 += t[2 * i]; m1 += t[2 * i + 1]; }
    m0 /= n; m1 /= n;
    double v0 = 0, v1 = 0;
    for (int i = 0; i < n; i++) {
        v0 += (t[2 * i] - m0) * (t[2 * i] - m0);
        v1 += (t[2 * i + 1] - m1) * (t[2 * i + 1] - m1);
    }
    double s0 = sqrt(v0 / (n - 1)), s1 = sqrt(v1 / (n - 1));
    m[0] = m0; m[1] = m1; s[0] = s0; s[1] = s1;
    if (!w) return;
    double sum = 0;
    for (int i = 0; i < n; i++) {
        double x = (t[2 * i] - m0) / s0, y = (t[2 * i + 1] - m1) / s1;
        double pdf = exp(-0.5 * (x * x + y * y)) / (2 * M_PI * s0 * s1);
        w[i] = 1.0 / pdf;
        sum += w[i];
    }
    for (int i = 0; i < n; i++) w[i] /= sum;
}

/*
 * Particle::setPeakFactor(PAR_R), 3-D (src/Particle.cpp:1893-1907):
 * f = clamp(u_(n/8) / u_(0), 1e-3, 0.5) over u sorted in descending order
 * (PEAK_FACTOR_BASE^3 = 8, include/Particle.h:44-49), when setFactor != 0;
 * then keepHalfHeightPeak (:1935-1954): hh = max(u) f, u_i <- u_i < hh ? 0 :
 * u_i - hh.  Returns f and the index of the first maximum (calRank1st).
 */
double thxo_keep_peak(double* u, int n, double factor, int setFactor, int* top)
{
    int best = 0;
    for (int i = 1; i < n; i++) if (u[i] > u[best]) best = i;
    if (setFactor) {
        double* s = (double*)malloc(sizeof(double) * n);
        memcpy(s, u, sizeof(double) * n);
        /* descending selection of rank n/8 by partial sort */
        int k = n / 8;
        for (int a = 0; a <= k; a++) {
            int m = a;
            for (int b = a + 1; b < n; b++) if (s[b] > s[m]) m = b;
            double tt = s[a]; s[a] = s[m]; s[m] = tt;
        }
        double f = s[k] / s[0];
        free(s);
        factor = f > 0.5 ? 0.5 : (f < 1e-3 ? 1e-3 : f);
    }
    double hh = u[best] * factor;
    for (int i = 0; i < n; i++) u[i] = u[i] < hh ? 0.0 : u[i] - hh;
    if (top) *top = best;
    return factor;
}

/*
 * a9: the post-scan update of one particle after the global scan of nCls
 * classes (src/Optimiser.cpp:926-1084; 3-D, OPTIMISER_PEAK_FACTOR_C and _R on,
 * _T off, PARTICLE_PEAK_FACTOR_C on, OPTIMISER_SCAN_SET_MIN_STD_WITH_PERTURB
 * on, PARTICLE_RHO off, include/Config.h:81-89, 214-224).  The random draws of
 * the reference (gsl_ran_flat for each systematic offset, gsl_rng_uniform_int
 * for rand(cls)) are given: u0C in [0, 1/nCls), pickC in [0, 1), u0R in
 * [0, 1/mLR), u0T in [0, 1/mLT); the shuffle before each resample is not
 * applied.
 *
 *   class: u_k = wC_k exp(base_k - max base) (setUC of the marginals on one
 *     baseline, :928-929); setPeakFactor(PAR_C) = PEAK_FACTOR_C = 1 - 1e-2
 *     (src/Particle.cpp:1879-1886, include/Particle.h:47) and
 *     keepHalfHeightPeak(PAR_C) (:1935-1944); resample(nCls, PAR_C) from the
 *     uniform class prior (:1280-1322: cdf of w u normalised by its last entry,
 *     c_j = first class with cdf >= u0C + j / nCls); rand(cls) = the draw
 *     floor(pickC nCls) (:2080-2089);
 *   rotations: setUR = wR of that class (:959-960), setPeakFactor(PAR_R)
 *     (order(nR / 8), clamped to [1e-3, 0.5]) + keepHalfHeightPeak(PAR_R)
 *     (:967-970), resample(mLR, PAR_R) with the grid prior pR (:1001;
 *     src/Particle.cpp:1324-1366), calVari(PAR_R) (:1004);
 *   translations: setUT = wT of that class (:961-962), no peak keeping,
 *     resample(mLT, PAR_T) with the grid prior pT (:1002), calVari(PAR_T)
 *     (:1005, n - 1 deviations, rho 0);
 *   clamps (:1031-1084): k_a = max(k_a, (minStdR / pfS)^2),
 *     s_a = max(s_a, minStdT / pfS), pfS = perturbFactorSGlobal.
 *
 * Inputs per class k: wC[k], base[k], wR[k R + r], wT[k T + t]; grid pR[R],
 * pT[T], gridQ[R][4], gridT[T][2].  Outputs: *cls, q[mLR][4] and wRn[mLR]
 * (resampled rotations and their 1/u weights, normalised), t[mLT][2], wTn[mLT],
 * k[3], s[2], *peakR (the rotation peak factor), idxR[mLR] / idxT[mLT] (grid
 * indices).  Returns 0, or the negative status of calVari.
 */
int thxo_post_scan(int nCls, const double* wC, const double* base, const double* wR,
                   const double* wT, const double* pR, const double* pT, const double* gridQ,
                   const double* gridT, int R, int T, int mLR, int mLT, double u0C,
                   double pickC, double u0R, double u0T, double minStdR, double minStdT,
                   double pfS, double tol, int maxIter, int* cls, int* idxR, double* q,
                   double* wRn, int* idxT, double* t, double* wTn, double* k, double* s,
                   double* peakR)
{
    /* class */
    int c = 0;
    if (nCls > 1) {
        double bmax = base[0], umax = 0, sum = 0;
        for (int i = 1; i < nCls; i++) if (base[i] > bmax) bmax = base[i];
        double* u = (double*)malloc(sizeof(double) * nCls);
        double* cdf = (double*)malloc(sizeof(double) * nCls);
        for (int i = 0; i < nCls; i++) {
            u[i] = wC[i] * exp(base[i] - bmax);
            if (u[i] > umax) umax = u[i];
        }
        const double hh = umax * (1 - 1e-2);
        for (int i = 0; i < nCls; i++) {
            u[i] = u[i] < hh ? 0.0 : u[i] - hh;
            sum += u[i];
        }
        /* w = uniform prior (.) u, normalised (all-zero u, which the peak
         * keeping cannot leave unless every wC is 0: the uniform prior) */
        double acc = 0;
        for (int i = 0; i < nCls; i++) {
            acc += sum > 0 ? u[i] / sum : 1.0 / nCls;
            cdf[i] = acc;
        }
        for (int i = 0; i < nCls; i++) cdf[i] /= cdf[nCls - 1];
        int j = (int)floor(pickC * nCls);
        if (j > nCls - 1) j = nCls - 1;
        const double uj = u0C + j * 1.0 / nCls;
        int i = 0;
        while (uj > cdf[i] && i < nCls - 1) i++;
        c = i;
        free(u);
        free(cdf);
    }
    *cls = c;
    /* rotations */
    double* u = (double*)malloc(sizeof(double) * (R > T ? R : T));
    memcpy(u, wR + (size_t)c * R, sizeof(double) * R);
    *peakR = thxo_keep_peak(u, R, 0.0, 1, NULL);
    thxo_resample(pR, u, R, mLR, u0R, idxR, wRn);
    for (int j = 0; j < mLR; j++) memcpy(q + 4 * j, gridQ + 4 * idxR[j], sizeof(double) * 4);
    /* translations */
    memcpy(u, wT + (size_t)c * T, sizeof(double) * T);
    thxo_resample(pT, u, T, mLT, u0T, idxT, wTn);
    for (int j = 0; j < mLT; j++) memcpy(t + 2 * j, gridT + 2 * idxT[j], sizeof(double) * 2);
    free(u);
    /* calVari and the scanning minima */
    double mean[4], m[2];
    int st = thxo_cal_vari_rot(q, mLR, tol, maxIter, mean, k);
    if (st < 0) return st;
    thxo_trans_stats(t, mLT, m, s, NULL);
    const double kMin = (minStdR / pfS) * (minStdR / pfS), sMin = minStdT / pfS;
    for (int a = 0; a < 3; a++) k[a] = k[a] > kMin ? k[a] : kMin;
    for (int a = 0; a < 2; a++) s[a] = s[a] > sMin ? s[a] : sMin;
    return 0;
}

/*
 * a13: the convergence test of one particle's local search
 * (src/Optimiser.cpp:1480-1612 with OPTIMISER_COMPRESS_CRITERIA,
 * include/Config.h:222; N_PHASE_WITH_NO_VARI_DECREASE = 1,
 * PARTICLE_FILTER_DECREASE_FACTOR = 0.95, include/Optimiser.h:56-63).
 * Phase i of the sequence is phase phase0 + i (phase0 = 1 after a global scan,
 * else 0); from phase minPhase on (10 after a global scan, else 3) the
 * particle's variances after that phase's update, variR = (k1 k2 k3)^(1/6)
 * (Particle::variR, src/Particle.cpp:609-625), variT = sqrt of the product of
 * the eigenvalues of [[s0^2, rho], [rho, s1^2]] (variT, :627-638) and
 * variD = sD (variD, :640-643), are compared with the smallest seen so far
 * (all DBL_MAX at first): a phase with none below 0.95 times its best ends the
 * search; the bests are then lowered.  k[3 i + a], s[3 i + a] (s0, s1, rho),
 * sD[i] for i < n.  Returns the phase that ended the search, or -1 when it
 * runs past the sequence.
 */
int thxo_converge(const double* k, const double* s, const double* sD, int n, int phase0,
                  int minPhase)
{
    double bR = DBL_MAX, bT = DBL_MAX, bD = DBL_MAX;
    int nNo = 0;
    for (int i = 0; i < n; i++) {
        const int phase = phase0 + i;
        if (phase < minPhase) continue;
        const double vR = pow(k[3 * i] * k[3 * i + 1] * k[3 * i + 2], 1.0 / 6);
        const double s0 = s[3 * i], s1 = s[3 * i + 1], rho = s[3 * i + 2];
        /* eigenvalues of the symmetric 2x2: their product is the determinant */
        const double vT = sqrt(s0 * s0 * s1 * s1 - rho * rho);
        const double vD = sD[i];
        if (vR < bR * 0.95 || vT < bT * 0.95 || vD < bD * 0.95) nNo = 0;
        else nNo += 1;
        if (vR < bR) bR = vR;
        if (vT < bT) bT = vT;
        if (vD < bD) bD = vD;
        if (nNo == 1) return phase;
    }
    return -1;
}

/*
 * F2: Reconstructor::insertP(const Complex* src, const RFLOAT* ctf,
 * const dmat33& rot, RFLOAT w, ...) (src/Reconstructor.cpp:782-860) with the
 * trilinear kernel and T added during the insert (RECONSTRUCTOR_TRILINEAR_KERNEL,
 * RECONSTRUCTOR_ADD_T_DURING_INSERT, include/Config.h:97-99), in FP64: for
 * each pixel k at padded coordinates (iCol pf, iRow pf, 0) rotated by M, the
 * value src_k ctf_k w goes to F and ctf_k^2 w to T through
 * Volume::addFT(value, x, y, z) (src/Image/Volume.cpp:334-368): mirrored and
 * conjugated (F only) when x < 0 (conjHalf), x0 = floor, weights
 * w[k][j][i] of WG_TRI_INTERP_LINEAR, corners wrapped in y and z
 * (addFTHalf, :555-700).  F[(z V + y)(V/2+1) + x] complex (re, im), T real.
 */
void thxo_insert(double* F, double* T, int V, int pf, const float* src, const float* ctf,
                 const double* M, const int* iCol, const int* iRow, int K, double w)
{
    const int nColFT = V / 2 + 1;
    for (int k = 0; k < K; k++) {
        const double c0 = (double)(iCol[k] * pf), c1 = (double)(iRow[k] * pf);
        double x = M[0] * c0 + M[1] * c1;
        double y = M[3] * c0 + M[4] * c1;
        double z = M[6] * c0 + M[7] * c1;
        double re = (double)src[2 * k] * ctf[k] * w, im = (double)src[2 * k + 1] * ctf[k] * w;
        const double t = (double)ctf[k] * ctf[k] * w;
        if (x < 0) { x = -x; y = -y; z = -z; im = -im; }
        const int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
        const double xd = x - x0, yd = y - y0, zd = z - z0;
        const double wx[2] = {1 - xd, xd}, wy[2] = {1 - yd, yd}, wz[2] = {1 - zd, zd};
        for (int a = 0; a < 2; a++) {
            int zz = z0 + a;
            zz = zz < 0 ? zz + V : zz;
            for (int b = 0; b < 2; b++) {
                int yy = y0 + b;
                yy = yy < 0 ? yy + V : yy;
                for (int c = 0; c < 2; c++) {
                    const double ww = wx[c] * wy[b] * wz[a];
                    const size_t idx = ((size_t)zz * V + yy) * nColFT + x0 + c;
                    F[2 * idx] += re * ww;
                    F[2 * idx + 1] += im * ww;
                    T[idx] += t * ww;
                }
            }
        }
    }
}
