// Particle filter of the local search on the GPU: the per-particle update
// around each local-search scoring phase, batched over all particles with
// uniform sample counts (mLR rotations, mLT translations per particle, as
// the reference's _para.mLR / _para.mLT).
//
// One phase (Optimiser::updateSP, src/Optimiser.cpp:1139-1437, 3-D, no CTF
// search) is: perturb rotations about the inferACG mean with ACG(1, pf^2
// min(1, k)) and rebalance their weights to 1 / pdfACG; perturb shifts with
// N(0, diag(s0^2, s1^2)) pf, re-centre outliers and rebalance to 1 / N2 pdf;
// score every (rotation, shift) sample pair (the local scorer); keep the
// half-height peak of the rotation marginal; record the top samples; infer
// the new concentrations k1..k3 (calVari) and shift deviations; resample
// both sets; and test convergence (src/Optimiser.cpp:1475-1660).
//
// Every stage kernel is one block (or one thread) per particle and skips the
// particles already converged.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <vector>

#include "thx_internal.hpp"

namespace thx {

// 4x4 inverse and determinant by the adjugate, from the twelve 2x2 minors of
// the upper and lower row pairs (Laplace expansion along those pairs; FP64).
// Returns false when the determinant is zero.
__device__ bool inv4(const double* A, double* inv, double* det) {
    const double a00 = A[0], a01 = A[1], a02 = A[2], a03 = A[3];
    const double a10 = A[4], a11 = A[5], a12 = A[6], a13 = A[7];
    const double a20 = A[8], a21 = A[9], a22 = A[10], a23 = A[11];
    const double a30 = A[12], a31 = A[13], a32 = A[14], a33 = A[15];
    const double s0 = a00 * a11 - a10 * a01, s1 = a00 * a12 - a10 * a02;
    const double s2 = a00 * a13 - a10 * a03, s3 = a01 * a12 - a11 * a02;
    const double s4 = a01 * a13 - a11 * a03, s5 = a02 * a13 - a12 * a03;
    const double c5 = a22 * a33 - a32 * a23, c4 = a21 * a33 - a31 * a23;
    const double c3 = a21 * a32 - a31 * a22, c2 = a20 * a33 - a30 * a23;
    const double c1 = a20 * a32 - a30 * a22, c0 = a20 * a31 - a30 * a21;
    const double d = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det) *det = d;
    if (d == 0.0) return false;
    const double r = 1.0 / d;
    inv[0] = (a11 * c5 - a12 * c4 + a13 * c3) * r;
    inv[1] = (-a01 * c5 + a02 * c4 - a03 * c3) * r;
    inv[2] = (a31 * s5 - a32 * s4 + a33 * s3) * r;
    inv[3] = (-a21 * s5 + a22 * s4 - a23 * s3) * r;
    inv[4] = (-a10 * c5 + a12 * c2 - a13 * c1) * r;
    inv[5] = (a00 * c5 - a02 * c2 + a03 * c1) * r;
    inv[6] = (-a30 * s5 + a32 * s2 - a33 * s1) * r;
    inv[7] = (a20 * s5 - a22 * s2 + a23 * s1) * r;
    inv[8] = (a10 * c4 - a11 * c2 + a13 * c0) * r;
    inv[9] = (-a00 * c4 + a01 * c2 - a03 * c0) * r;
    inv[10] = (a30 * s4 - a31 * s2 + a33 * s0) * r;
    inv[11] = (-a20 * s4 + a21 * s2 - a23 * s0) * r;
    inv[12] = (-a10 * c3 + a11 * c1 - a12 * c0) * r;
    inv[13] = (a00 * c3 - a01 * c1 + a02 * c0) * r;
    inv[14] = (-a30 * s3 + a31 * s1 - a32 * s0) * r;
    inv[15] = (a20 * s3 - a21 * s1 + a22 * s0) * r;
    return true;
}

// Eigenvector of the largest eigenvalue of a symmetric 4x4 (cyclic Jacobi),
// unit length, largest-magnitude entry positive.
__device__ void sym4_top(const double* Ain, double* vec) {
    double A[4][4], V[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            A[i][j] = Ain[4 * i + j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) off += A[a][b] * A[a][b];
        if (off < 1e-40) break;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) {
                if (fabs(A[a][b]) < 1e-300) continue;
                const double th = (A[b][b] - A[a][a]) / (2 * A[a][b]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; k++) {
                    const double x = A[k][a], y = A[k][b];
                    A[k][a] = c * x - s * y;
                    A[k][b] = s * x + c * y;
                }
                for (int k = 0; k < 4; k++) {
                    const double x = A[a][k], y = A[b][k];
                    A[a][k] = c * x - s * y;
                    A[b][k] = s * x + c * y;
                }
                for (int k = 0; k < 4; k++) {
                    const double x = V[k][a], y = V[k][b];
                    V[k][a] = c * x - s * y;
                    V[k][b] = s * x + c * y;
                }
            }
    }
    int top = 0;
    for (int a = 1; a < 4; a++)
        if (A[a][a] > A[top][top]) top = a;
    int big = 0;
    for (int a = 1; a < 4; a++)
        if (fabs(V[a][top]) > fabs(V[big][top])) big = a;
    const double sg = V[big][top] < 0 ? -1.0 : 1.0;
    double n = 0.0;
    for (int a = 0; a < 4; a++) n += V[a][top] * V[a][top];
    n = sqrt(n);
    for (int a = 0; a < 4; a++) vec[a] = sg * V[a][top] / n;
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide sum (256 threads); every thread gets the total.
__device__ double block_sum(double v, double* part) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return part[0] + part[1] + part[2] + part[3];
}

// Fixed point of inferACG(dmat44&, const dmat4&) (src/Geometry/
// DirectionalStat.cpp:87-139) over n samples, each optionally left-multiplied
// by the quaternion pre: A <- I; repeat B = (4 / sum 1/u) sum x x^T / u,
// u = x^T A^-1 x, until sum |A - B| <= tol; the result is the last A.
// A, Ai (shared, 16 each), part (shared, 44), flag (shared).  Returns the
// iteration count (negative when maxIter stopped it or A became singular).
__device__ int block_infer_acg(const double* __restrict__ q, int n, const double* pre,
                               double tol, int maxIter, double* A, double* Ai, double* part,
                               int* flag) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 16) A[tid] = (tid % 5 == 0) ? 1.0 : 0.0;
    __syncthreads();
    int it = 0, status = 0;
    for (;;) {
        if (tid == 0) *flag = inv4(A, Ai, nullptr) ? 0 : 2;
        __syncthreads();
        if (*flag == 2) { status = -1; break; }
        double acc[11];
        for (int j = 0; j < 11; j++) acc[j] = 0.0;
        for (int i = tid; i < n; i += blockDim.x) {
            double x[4] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
            if (pre) {
                double y[4];
                qmul(pre, x, y);
                for (int a = 0; a < 4; a++) x[a] = y[a];
            }
            double u = 0.0;
            for (int a = 0; a < 4; a++) {
                double v = 0.0;
                for (int b = 0; b < 4; b++) v += Ai[4 * a + b] * x[b];
                u += x[a] * v;
            }
            const double iu = 1.0 / u;
            int j = 0;
            for (int a = 0; a < 4; a++)
                for (int b = a; b < 4; b++) acc[j++] += x[a] * x[b] * iu;
            acc[10] += iu;
        }
        for (int j = 0; j < 11; j++) {
            const double v = wave_sum(acc[j]);
            if (lane == 0) part[wave * 11 + j] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double sm[11];
            for (int j = 0; j < 11; j++)
                sm[j] = part[j] + part[11 + j] + part[22 + j] + part[33 + j];
            const double f = 4.0 / sm[10];
            double B[16];
            int j = 0;
            for (int a = 0; a < 4; a++)
                for (int b = a; b < 4; b++) {
                    B[4 * a + b] = B[4 * b + a] = f * sm[j];
                    j++;
                }
            double diff = 0.0;
            for (int e = 0; e < 16; e++) diff += fabs(A[e] - B[e]);
            if (diff <= tol) {
                *flag = 1;
            } else if (it + 1 >= maxIter) {
                *flag = 3;
            } else {
                for (int e = 0; e < 16; e++) A[e] = B[e];
                *flag = 0;
            }
        }
        __syncthreads();
        it++;
        if (*flag == 1) break;
        if (*flag == 3) { status = -2; break; }
    }
    __syncthreads();
    return status < 0 ? status : it;
}

// Per particle over its n rotation samples (flags): 1 = balance weights
// w_i = 1 / pdfACG(q_i, A) normalised (Particle::balanceWeight(PAR_R),
// src/Particle.cpp:2301-2311; pdfACG = det(A)^-1/2 (x^T A^-1 x)^-2,
// src/Geometry/DirectionalStat.cpp:13-18); 2 = mean (inferACG(dvec4&, ...),
// :218-244); 4 = calVari concentrations k_a = A'(a,a) / A'(0,0) of inferACG
// of conj(mean) q (src/Particle.cpp:1043-1067, DirectionalStat.cpp:178-216).
// 8 = also write the balance weights as FP32 priors for the scorer.
__global__ __launch_bounds__(256) void k_rot_stats(
    const double* __restrict__ quat, int n, const int* __restrict__ active, int flags, double tol,
    int maxIter, double* __restrict__ Aout, double* __restrict__ meanOut,
    double* __restrict__ kOut, double* __restrict__ wBal, float* __restrict__ wBalF,
    int* __restrict__ itOut) {
    const int p = blockIdx.x;
    if (active && !active[p]) return;
    __shared__ double A[16], Ai[16], part[44], pre[4], det;
    __shared__ int flag;
    const double* q = quat + (size_t)p * n * 4;
    int it = block_infer_acg(q, n, nullptr, tol, maxIter, A, Ai, part, &flag);
    if (Aout && threadIdx.x < 16) Aout[(size_t)p * 16 + threadIdx.x] = A[threadIdx.x];
    if (flags & 1) {
        if (threadIdx.x == 0) {
            double d = 1.0;
            if (!inv4(A, Ai, &d)) d = 0.0;
            det = d;
        }
        __syncthreads();
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double* x = q + 4 * i;
            double u = 0.0;
            for (int a = 0; a < 4; a++) {
                double v = 0.0;
                for (int b = 0; b < 4; b++) v += Ai[4 * a + b] * x[b];
                u += x[a] * v;
            }
            s += sqrt(fmax(det, 0.0)) * u * u;
        }
        const double tot = block_sum(s, part);
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double* x = q + 4 * i;
            double u = 0.0;
            for (int a = 0; a < 4; a++) {
                double v = 0.0;
                for (int b = 0; b < 4; b++) v += Ai[4 * a + b] * x[b];
                u += x[a] * v;
            }
            const double w = sqrt(fmax(det, 0.0)) * u * u / tot;
            wBal[(size_t)p * n + i] = w;
            if (flags & 8) wBalF[(size_t)p * n + i] = (float)w;
        }
        __syncthreads();
    }
    if (flags & 6) {
        if (threadIdx.x == 0) {
            double m[4];
            sym4_top(A, m);
            pre[0] = m[0];
            pre[1] = -m[1];
            pre[2] = -m[2];
            pre[3] = -m[3];
            if (meanOut)
                for (int a = 0; a < 4; a++) meanOut[4 * (size_t)p + a] = m[a];
        }
        __syncthreads();
        if (flags & 4) {
            const int it2 = block_infer_acg(q, n, pre, tol, maxIter, A, Ai, part, &flag);
            if (threadIdx.x == 0)
                for (int a = 0; a < 3; a++) kOut[3 * (size_t)p + a] = A[5 * (a + 1)] / A[0];
            if (it2 < 0) it = it2;
        }
    }
    if (itOut && threadIdx.x == 0) itOut[p] = it;
}

// Shifts of one particle (one thread per particle, n samples): optional
// perturbation t += pf (s0 g0, s1 g1) (Particle::perturb(pf, PAR_T),
// src/Particle.cpp:1227-1262, rho = 0) with re-centring of samples farther
// than transM from the origin to N(0, transS^2 I) (reCentre, :2427-2449);
// then means, n - 1 standard deviations (calVari(PAR_T), :1083-1097) and the
// balance weights 1 / N2(t - m; s0, s1) normalised (balanceWeight(PAR_T),
// :2313-2344).  s01 (in: the perturbation's deviations; out: the new ones).
__global__ void k_trans_stats(double* __restrict__ trans, int n, const int* __restrict__ active,
                              int P, int perturb, double* __restrict__ s01, PhaseScale pfs,
                              const double* __restrict__ transS, double transQ, uint64_t seed,
                              PhaseArg pha, double* __restrict__ mOut, double* __restrict__ wBal,
                              float* __restrict__ wBalF) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || (active && !active[p])) return;
    const double pf = pfs.get();
    const uint32_t phase = pha.get();
    double* t = trans + (size_t)p * n * 2;
    if (perturb) {
        const double s0 = s01[3 * p], s1 = s01[3 * p + 1];
        const double tS = transS[p];
        const double transM = tS * (-2.0 * log(transQ));   // tS * chi2_2 Qinv(transQ)
        for (int i = 0; i < n; i++) {
            double g[4];
            normals4(seed, (uint32_t)((size_t)p * n + i), phase, kStreamShift + 48, g);
            t[2 * i] += pf * s0 * g[0];
            t[2 * i + 1] += pf * s1 * g[1];
            if (sqrt(t[2 * i] * t[2 * i] + t[2 * i + 1] * t[2 * i + 1]) > transM) {
                t[2 * i] = tS * g[2];
                t[2 * i + 1] = tS * g[3];
            }
        }
    }
    double m0 = 0.0, m1 = 0.0;
    for (int i = 0; i < n; i++) {
        m0 += t[2 * i];
        m1 += t[2 * i + 1];
    }
    m0 /= n;
    m1 /= n;
    double v0 = 0.0, v1 = 0.0;
    for (int i = 0; i < n; i++) {
        v0 += (t[2 * i] - m0) * (t[2 * i] - m0);
        v1 += (t[2 * i + 1] - m1) * (t[2 * i + 1] - m1);
    }
    const double s0 = sqrt(v0 / (n - 1)), s1 = sqrt(v1 / (n - 1));
    s01[3 * p] = s0;
    s01[3 * p + 1] = s1;
    s01[3 * p + 2] = 0.0;
    if (mOut) {
        mOut[2 * p] = m0;
        mOut[2 * p + 1] = m1;
    }
    if (!wBal) return;
    // 1 / pdf up to the common factor 2 pi s0 s1, normalised
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
        const double x = (t[2 * i] - m0) / s0, y = (t[2 * i + 1] - m1) / s1;
        sum += exp(0.5 * (x * x + y * y));
    }
    for (int i = 0; i < n; i++) {
        const double x = (t[2 * i] - m0) / s0, y = (t[2 * i + 1] - m1) / s1;
        const double w = exp(0.5 * (x * x + y * y)) / sum;
        wBal[(size_t)p * n + i] = w;
        if (wBalF) wBalF[(size_t)p * n + i] = (float)w;
    }
}

// setPeakFactor (setFactor) + keepHalfHeightPeak + calRank1st of one sample
// type (src/Particle.cpp:988-1000, 1864-1954): u (n per particle, >= 0) in
// place.  The factor is clamp(u_(n/8) / u_(0), 1e-3, 0.5) over the descending
// order (rotations, 3-D); hh = max(u) factor; u_i <- u_i < hh ? 0 : u_i - hh.
// topOut receives the sample (dim values) of the first maximum.  One block per
// particle; the rank statistic is found by bisection on the FP64 bit pattern.
__global__ __launch_bounds__(256) void k_keep_peak(
    double* __restrict__ u, int n, const int* __restrict__ active, double* __restrict__ factor,
    int setFactor, int keep, const double* __restrict__ samples, int dim,
    double* __restrict__ topOut, int* __restrict__ topIdx) {
    const int p = blockIdx.x;
    if (active && !active[p]) return;
    double* up = u + (size_t)p * n;
    __shared__ double rmax[256];
    __shared__ int ridx[256];
    __shared__ double part[4];
    __shared__ double sh[2];
    double mx = -1.0;
    int mi = 0x7fffffff;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (up[i] > mx) { mx = up[i]; mi = i; }
    rmax[threadIdx.x] = mx;
    ridx[threadIdx.x] = mi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = rmax[threadIdx.x + s];
            const int oi = ridx[threadIdx.x + s];
            if (o > rmax[threadIdx.x] || (o == rmax[threadIdx.x] && oi < ridx[threadIdx.x])) {
                rmax[threadIdx.x] = o;
                ridx[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double umax = rmax[0];
    const int best = ridx[0];
    if (setFactor) {
        // largest key K with count(key >= K) >= n/8 + 1 (keys of u >= 0 order as u)
        const unsigned long long need = (unsigned long long)(n / 8 + 1);
        unsigned long long lo = 0ull, hi = (unsigned long long)__double_as_longlong(umax);
        while (lo < hi) {
            const unsigned long long mid = lo + (hi - lo + 1) / 2;
            double c = 0.0;
            for (int i = threadIdx.x; i < n; i += blockDim.x)
                c += (unsigned long long)__double_as_longlong(up[i]) >= mid ? 1.0 : 0.0;
            const double tot = block_sum(c, part);
            if (tot >= (double)need) lo = mid; else hi = mid - 1;
        }
        if (threadIdx.x == 0) {
            const double f = umax > 0 ? __longlong_as_double((long long)lo) / umax : 0.0;
            factor[p] = fmin(0.5, fmax(1e-3, f));
        }
        __syncthreads();
    }
    if (keep) {
        if (threadIdx.x == 0) sh[0] = umax * factor[p];
        __syncthreads();
        const double hh = sh[0];
        for (int i = threadIdx.x; i < n; i += blockDim.x) up[i] = up[i] < hh ? 0.0 : up[i] - hh;
    }
    if (threadIdx.x < dim && topOut)
        topOut[(size_t)p * dim + threadIdx.x] = samples[((size_t)p * n + best) * dim + threadIdx.x];
    if (threadIdx.x == 0 && topIdx) topIdx[p] = best;
}

// After systematic resampling: particle p's new samples are its (or the
// shared grid's, srcShared) samples idx[p][j]; the new weights wNew are its
// prior (and its FP32 copy for the scorer).  Converged particles are left.
__global__ void k_gather_samples(const double* __restrict__ src, int nSrc, int srcShared,
                                 const int* __restrict__ idx, const double* __restrict__ wNew,
                                 int n, int dim, const int* __restrict__ active, int P,
                                 double* __restrict__ dst, double* __restrict__ w,
                                 float* __restrict__ wF) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)P * n) return;
    const int p = (int)(g / n);
    if (active && !active[p]) return;
    const size_t s = (srcShared ? 0 : (size_t)p * nSrc) + idx[g];
    for (int d = 0; d < dim; d++) dst[g * dim + d] = src[s * dim + d];
    w[g] = wNew[g];
    if (wF) wF[g] = (float)wNew[g];
}

// Rows of converged particles copied into the other buffer of a pair.
__global__ void k_keep_rows(const double* __restrict__ src, double* __restrict__ dst, int rowLen,
                            const int* __restrict__ active, int P) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)P * rowLen) return;
    if (active[g / rowLen]) return;
    dst[g] = src[g];
}

// Convergence of the local search (src/Optimiser.cpp:1511-1620 with
// OPTIMISER_COMPRESS_CRITERIA, include/Config.h:222): variR = (k1 k2 k3)^(1/6)
// (Particle::variR, src/Particle.cpp:609-625), variT = s0 s1 (variT, :627-638,
// rho = 0).  From phase minPhase on, a phase with neither below 0.95 x the
// best so far counts; the first such phase ends the particle's search
// (N_PHASE_WITH_NO_VARI_DECREASE = 1, include/Optimiser.h:56-63).
__global__ void k_converge(const double* __restrict__ kk, const double* __restrict__ s01,
                           int* __restrict__ active, double* __restrict__ best,
                           int* __restrict__ nPhase, int P, PhaseArg pha, int minPhase) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !active[p]) return;
    const int phase = (int)pha.get();
    if (phase < minPhase) return;
    const double vR = pow(kk[3 * p] * kk[3 * p + 1] * kk[3 * p + 2], 1.0 / 6);
    const double vT = s01[3 * p] * s01[3 * p + 1];
    const double bR = best[2 * p], bT = best[2 * p + 1];
    // the defocus term of the reference (no CTF search) only counts on the
    // first check, where every best is still DBL_MAX
    const bool first = bR == 1.7976931348623157e308;
    const bool dec = first || vR < bR * 0.95 || vT < bT * 0.95;
    if (vR < bR) best[2 * p] = vR;
    if (vT < bT) best[2 * p + 1] = vT;
    if (!dec) {
        active[p] = 0;
        nPhase[p] = phase;
    }
}

// min-std clamps after the global scan (src/Optimiser.cpp:1031-1084 with
// OPTIMISER_SCAN_SET_MIN_STD_WITH_PERTURB): k_a >= (minR / pfS)^2,
// s_a >= minT / pfS.
__global__ void k_min_std(double* __restrict__ kk, double* __restrict__ s01, int P, double kMin,
                          double sMin) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int a = 0; a < 3; a++) kk[3 * p + a] = fmax(kMin, kk[3 * p + a]);
    for (int a = 0; a < 2; a++) s01[3 * p + a] = fmax(sMin, s01[3 * p + a]);
}

// The 4x4 helpers of the inferACG stages on n given matrices (checking):
// inverse and determinant (inv4), top eigenvector of the symmetric part
// (sym4_top).  One thread per matrix.
__global__ void k_mat4_check(const double* __restrict__ A, int n, double* __restrict__ inv,
                             double* __restrict__ det, double* __restrict__ top) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[16], b[16], d = 0.0;
    for (int e = 0; e < 16; e++) a[e] = A[16 * (size_t)i + e];
    if (!inv4(a, b, &d))
        for (int e = 0; e < 16; e++) b[e] = 0.0;
    for (int e = 0; e < 16; e++) inv[16 * (size_t)i + e] = b[e];
    det[i] = d;
    double v[4];
    sym4_top(a, v);
    for (int e = 0; e < 4; e++) top[4 * (size_t)i + e] = v[e];
}

constexpr double kAcgTol = 1e-3;     // inferACG's convergence bound
constexpr int kAcgMaxIter = 1000;
constexpr double kTransQ = 0.05;     // TRANS_Q, include/Optimiser.h:67

static unsigned nb(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace thx

using namespace thx;

namespace {

int pf_alloc(thx_ctx* c, int P, int mLR, int mLT) {
    c->pfP = P;
    c->pfR = mLR;
    c->pfT = mLT;
    const size_t nr = (size_t)P * mLR, nt = (size_t)P * mLT;
    THX_TRY(c->pfQ.ensure(sizeof(double) * 4 * nr));
    THX_TRY(c->pfQ2.ensure(sizeof(double) * 4 * nr));
    THX_TRY(c->pfTr.ensure(sizeof(double) * 2 * nt));
    THX_TRY(c->pfTr2.ensure(sizeof(double) * 2 * nt));
    THX_TRY(c->pfWR.ensure(sizeof(double) * nr));
    THX_TRY(c->pfWT.ensure(sizeof(double) * nt));
    THX_TRY(c->pfWRf.ensure(sizeof(float) * nr));
    THX_TRY(c->pfWTf.ensure(sizeof(float) * nt));
    THX_TRY(c->pfIdx.ensure(sizeof(int) * std::max(nr, nt)));
    THX_TRY(c->pfWNew.ensure(sizeof(double) * std::max(nr, nt)));
    THX_TRY(c->pfK.ensure(sizeof(double) * 3 * P));
    THX_TRY(c->pfS.ensure(sizeof(double) * 3 * P));
    THX_TRY(c->pfMean.ensure(sizeof(double) * 4 * P));
    THX_TRY(c->pfTopR.ensure(sizeof(double) * 4 * P));
    THX_TRY(c->pfTopT.ensure(sizeof(double) * 2 * P));
    THX_TRY(c->pfPeak.ensure(sizeof(double) * P));
    THX_TRY(c->pfTransS.ensure(sizeof(double) * P));
    THX_TRY(c->pfBest.ensure(sizeof(double) * 2 * P));
    THX_TRY(c->pfActive.ensure(sizeof(int) * P));
    THX_TRY(c->pfNPhase.ensure(sizeof(int) * P));
    std::vector<int> offR(P), offT(P), nR(P, mLR), nT(P, mLT);
    for (int p = 0; p < P; p++) {
        offR[p] = p * mLR;
        offT[p] = p * mLT;
    }
    THX_HIP(hipStreamSynchronize(c->stream));
    const std::vector<int> ones(P, 1);
    THX_HIP(hipMemcpy(c->pfActive.p, ones.data(), sizeof(int) * P, hipMemcpyHostToDevice));
    std::vector<double> best(2 * (size_t)P, 1.7976931348623157e308);
    std::vector<int> np(P, -1);
    THX_HIP(hipMemcpy(c->pfBest.p, best.data(), sizeof(double) * 2 * P, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pfNPhase.p, np.data(), sizeof(int) * P, hipMemcpyHostToDevice));
    auto put = [&](DevBuf& d, const std::vector<int>& v) -> int {
        THX_TRY(d.ensure(sizeof(int) * v.size()));
        THX_HIP(hipMemcpy(d.p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice));
        return THX_OK;
    };
    THX_TRY(put(c->pfOffR, offR));
    THX_TRY(put(c->pfOffT, offT));
    THX_TRY(put(c->pfNR, nR));
    THX_TRY(put(c->pfNT, nT));
    c->pfValid = true;
    return THX_OK;
}

}  // namespace

extern "C" {

int thx_pf_set_state(thx_ctx* c, int mLR, int mLT, const double* quat, const double* trans,
                     const double* k, const double* s01, const double* peakR,
                     const double* transS) {
    clear_error();
    THX_CHECK(c && quat && trans && k && s01 && peakR && transS, THX_ERR_ARG,
              "pf_set_state: null argument");
    THX_CHECK(c->P > 0, THX_ERR_STATE, "pf_set_state: particles must be set first");
    THX_CHECK(mLR >= 2 && mLR <= kResampleMaxN && mLT >= 2 && mLT <= 128, THX_ERR_RANGE,
              "pf_set_state: mLR must be in [2, 8192] and mLT in [2, 128]");
    THX_HIP(hipSetDevice(c->device));
    const int P = c->P;
    THX_TRY(pf_alloc(c, P, mLR, mLT));
    THX_HIP(hipMemcpy(c->pfQ.p, quat, sizeof(double) * 4 * P * (size_t)mLR, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pfTr.p, trans, sizeof(double) * 2 * P * (size_t)mLT, hipMemcpyHostToDevice));
    std::vector<double> s3(3 * (size_t)P);
    for (int p = 0; p < P; p++) {
        s3[3 * p] = s01[2 * p];
        s3[3 * p + 1] = s01[2 * p + 1];
        s3[3 * p + 2] = 0.0;
    }
    THX_HIP(hipMemcpy(c->pfK.p, k, sizeof(double) * 3 * P, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pfS.p, s3.data(), sizeof(double) * 3 * P, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pfPeak.p, peakR, sizeof(double) * P, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pfTransS.p, transS, sizeof(double) * P, hipMemcpyHostToDevice));
    return THX_OK;
}

// dst[p][0..n) = rows[cls[p]][p][0..n): each particle's marginal from its class
__global__ void k_gather_class_rows(const double* const* __restrict__ rows,
                                    const int* __restrict__ cls, int n,
                                    double* __restrict__ dst) {
    const int p = blockIdx.x;
    const double* src = rows[cls[p]] + (size_t)p * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[(size_t)p * n + i] = src[i];
}

static int check_scan(thx_ctx* c, int k) {
    THX_CHECK(k >= 0 && k < (int)c->res.size() && c->res[k].valid, THX_ERR_STATE,
              "pf_from_global: no global scan of this class");
    THX_CHECK(c->res[k].wR.bytes >= sizeof(double) * (size_t)c->P * c->R &&
                  c->res[k].wT.bytes >= sizeof(double) * (size_t)c->P * c->T,
              THX_ERR_STATE, "pf_from_global: marginals of the last scan do not match the grid");
    return THX_OK;
}

// Particle states from the rotation / translation marginals of class cls0
// (pcls == null) or of each particle's own class pcls[p] (host array).
// u0R / u0T (host, P each, or null: drawn from the seed) are the systematic
// resampling offsets of the rotations and translations.
static int from_global_impl(thx_ctx* c, int cls0, const int* pcls, int nCls, int mLR, int mLT,
                            double minStdR, double minStdT, double pfS, const double* transS,
                            uint64_t seed, const double* u0R = nullptr,
                            const double* u0T = nullptr) {
    THX_CHECK(mLR >= 2 && mLR <= kResampleMaxN && mLT >= 2 && mLT <= 128 &&
                  c->R <= kResampleMaxN && c->T <= kResampleMaxN,
              THX_ERR_RANGE, "pf_from_global: sizes out of range");
    THX_CHECK(c->pR.p && c->pT.p && c->pR.bytes >= sizeof(double) * c->R &&
                  c->pT.bytes >= sizeof(double) * c->T && c->quat.p && c->trans.p,
              THX_ERR_STATE, "pf_from_global: grid and its priors must be set");
    THX_HIP(hipSetDevice(c->device));
    const int P = c->P, R = c->R, T = c->T;
    hipStream_t st = c->stream;
    THX_TRY(pf_alloc(c, P, mLR, mLT));
    THX_HIP(hipMemcpy(c->pfTransS.p, transS, sizeof(double) * P, hipMemcpyHostToDevice));
    // marginals of the grid: copies (the peak keeping works in place)
    DevBuf uR, uT, offR, offT, nRg, nTg, dU0R, dU0T;
    std::vector<int> oR(P), oT(P), cR(P, R), cT(P, T);
    for (int p = 0; p < P; p++) {
        oR[p] = p * R;
        oT[p] = p * T;
    }
    int s = THX_OK;
    auto cleanup = [&]() {
        DevBuf* b[] = {&uR, &uT, &offR, &offT, &nRg, &nTg, &dU0R, &dU0T};
        for (DevBuf* x : b) x->release();
    };
    if ((s = uR.ensure(sizeof(double) * (size_t)P * R)) || (s = uT.ensure(sizeof(double) * (size_t)P * T)) ||
        (s = offR.ensure(sizeof(int) * P)) || (s = offT.ensure(sizeof(int) * P)) ||
        (s = nRg.ensure(sizeof(int) * P)) || (s = nTg.ensure(sizeof(int) * P)) ||
        (u0R && (s = upload(dU0R, u0R, P, st))) || (u0T && (s = upload(dU0T, u0T, P, st)))) {
        cleanup();
        return s;
    }
    if (!pcls) {
        const ClassResult& res = c->res[cls0];
        THX_HIP(hipMemcpyAsync(uR.p, res.wR.p, sizeof(double) * (size_t)P * R,
                               hipMemcpyDeviceToDevice, st));
        THX_HIP(hipMemcpyAsync(uT.p, res.wT.p, sizeof(double) * (size_t)P * T,
                               hipMemcpyDeviceToDevice, st));
    } else {
        std::vector<const double*> rowsR(nCls), rowsT(nCls);
        for (int k = 0; k < nCls; k++) {
            rowsR[k] = c->res[k].wR.as<const double>();
            rowsT[k] = c->res[k].wT.as<const double>();
        }
        DevBuf dR, dT, dC;
        int e = THX_OK;
        if ((e = dR.ensure(sizeof(void*) * nCls)) || (e = dT.ensure(sizeof(void*) * nCls)) ||
            (e = dC.ensure(sizeof(int) * P))) {
            cleanup();
            return e;
        }
        bool ok = hipMemcpy(dR.p, rowsR.data(), sizeof(void*) * nCls, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(dT.p, rowsT.data(), sizeof(void*) * nCls, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(dC.p, pcls, sizeof(int) * P, hipMemcpyHostToDevice) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(k_gather_class_rows, dim3(P), dim3(256), 0, st,
                               dR.as<const double* const>(), dC.as<const int>(), R, uR.as<double>());
            hipLaunchKernelGGL(k_gather_class_rows, dim3(P), dim3(256), 0, st,
                               dT.as<const double* const>(), dC.as<const int>(), T, uT.as<double>());
            ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        }
        dR.release();
        dT.release();
        dC.release();
        if (!ok) {
            cleanup();
            set_error("pf_from_global: gathering the class marginals failed");
            return THX_ERR_HIP;
        }
    }
    THX_HIP(hipMemcpyAsync(offR.p, oR.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
    THX_HIP(hipMemcpyAsync(offT.p, oT.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
    THX_HIP(hipMemcpyAsync(nRg.p, cR.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
    THX_HIP(hipMemcpyAsync(nTg.p, cT.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
    THX_HIP(hipStreamSynchronize(st));
    // setUR + setPeakFactor + keepHalfHeightPeak (rotations), rank of both
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, uR.as<double>(), R, nullptr,
                       c->pfPeak.as<double>(), 1, 1, nullptr, 4, nullptr, nullptr);
    THX_HIP(hipGetLastError());
    // resample mLR rotations / mLT shifts from the grid (prior = the grid's)
    s = pf_resample(st, c->pR.as<double>(), uR.as<double>(), offR.as<int>(), nRg.as<int>(),
                    c->pfOffR.as<int>(), c->pfNR.as<int>(), P, seed, 0u,
                    u0R ? dU0R.as<const double>() : nullptr, c->pfIdx.as<int>(),
                    c->pfWNew.as<double>(), nullptr, 1);
    if (s) { cleanup(); return s; }
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLR)), dim3(256), 0, st,
                       c->quat.as<const double>(), R, 1, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLR, 4, nullptr, P, c->pfQ.as<double>(),
                       c->pfWR.as<double>(), c->pfWRf.as<float>());
    s = pf_resample(st, c->pT.as<double>(), uT.as<double>(), offT.as<int>(), nTg.as<int>(),
                    c->pfOffT.as<int>(), c->pfNT.as<int>(), P, seed, 1u,
                    u0T ? dU0T.as<const double>() : nullptr, c->pfIdx.as<int>(),
                    c->pfWNew.as<double>(), nullptr, 1);
    if (s) { cleanup(); return s; }
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLT)), dim3(256), 0, st,
                       c->trans.as<const double>(), T, 1, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLT, 2, nullptr, P, c->pfTr.as<double>(),
                       c->pfWT.as<double>(), c->pfWTf.as<float>());
    // calVari of both sets, then the scanning minimum deviations
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, c->pfQ.as<const double>(), mLR,
                       nullptr, 6, kAcgTol, kAcgMaxIter, nullptr, c->pfMean.as<double>(),
                       c->pfK.as<double>(), nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_trans_stats, dim3(nb(P)), dim3(256), 0, st, c->pfTr.as<double>(), mLT,
                       nullptr, P, 0, c->pfS.as<double>(), scale_host(0.0),
                       c->pfTransS.as<const double>(), kTransQ, seed, phase_host(0u), nullptr,
                       nullptr, nullptr);
    const double kMin = (minStdR / pfS) * (minStdR / pfS), sMin = minStdT / pfS;
    hipLaunchKernelGGL(k_min_std, dim3(nb(P)), dim3(256), 0, st, c->pfK.as<double>(),
                       c->pfS.as<double>(), P, kMin, sMin);
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    cleanup();
    return THX_OK;
}

int thx_pf_from_global(thx_ctx* c, int cls, int mLR, int mLT, double minStdR, double minStdT,
                       double pfS, const double* transS, uint64_t seed) {
    clear_error();
    THX_CHECK(c && transS && pfS > 0, THX_ERR_ARG, "pf_from_global: bad argument");
    THX_TRY(check_scan(c, cls));
    return from_global_impl(c, cls, nullptr, 0, mLR, mLT, minStdR, minStdT, pfS, transS, seed);
}

// Class of every particle from the class scans 0..nCls-1 (src/Optimiser.cpp:
// 926-966): the class marginals on a common baseline, u_k = wC_k exp(base_k -
// max base), keep their half-height peak with the fixed factor 1 - 1e-2
// (setPeakFactor / keepHalfHeightPeak(PAR_C), src/Particle.cpp:1879-1886,
// 1935-1944; include/Particle.h:47), are resampled to nCls systematic draws
// u0 + j / nCls over the cumulative weights normalised by their last entry,
// each taking the first class whose cdf reaches it (resample(k, PAR_C),
// src/Particle.cpp:1280-1322), and the draw floor(pick nCls) is the class
// (rand(cls), :2080-2089).  u0[p] in [0, 1/nCls), pick[p] in [0, 1).
static int draw_classes(thx_ctx* c, int nCls, const double* u0, const double* pick, int* clsOut) {
    THX_HIP(hipSetDevice(c->device));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int P = c->P;
    std::vector<double> wC((size_t)nCls * P), base((size_t)nCls * P);
    for (int k = 0; k < nCls; k++) {
        THX_HIP(hipMemcpy(wC.data() + (size_t)k * P, c->res[k].wC.p, sizeof(double) * P,
                          hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(base.data() + (size_t)k * P, c->res[k].base.p, sizeof(double) * P,
                          hipMemcpyDeviceToHost));
    }
    std::vector<double> u(nCls), cdf(nCls);
    for (int p = 0; p < P; p++) {
        double bmax = -INFINITY;
        for (int k = 0; k < nCls; k++) bmax = std::max(bmax, base[(size_t)k * P + p]);
        double umax = 0.0;
        for (int k = 0; k < nCls; k++) {
            u[k] = wC[(size_t)k * P + p] * std::exp(base[(size_t)k * P + p] - bmax);
            umax = std::max(umax, u[k]);
        }
        const double hh = umax * (1.0 - 1e-2);
        double sum = 0.0;
        for (int k = 0; k < nCls; k++) {
            u[k] = u[k] < hh ? 0.0 : u[k] - hh;
            sum += u[k];
        }
        // uniform prior times u, normalised (all-zero u: the uniform prior)
        double acc = 0.0;
        for (int k = 0; k < nCls; k++) {
            acc += sum > 0.0 ? u[k] / sum : 1.0 / nCls;
            cdf[k] = acc;
        }
        const double last = cdf[nCls - 1];
        for (int k = 0; k < nCls; k++) cdf[k] /= last;
        const int j = std::min(nCls - 1, (int)std::floor(pick[p] * nCls));
        const double uj = u0[p] + (double)j / nCls;
        int k = 0;
        while (k < nCls - 1 && uj > cdf[k]) k++;
        clsOut[p] = k;
    }
    return THX_OK;
}

int thx_pf_from_global_classes(thx_ctx* c, int nCls, int mLR, int mLT, double minStdR,
                               double minStdT, double pfS, const double* transS, uint64_t seed,
                               int* clsOut) {
    clear_error();
    THX_CHECK(c && transS && clsOut && pfS > 0 && nCls >= 1, THX_ERR_ARG,
              "pf_from_global_classes: bad argument");
    for (int k = 0; k < nCls; k++) THX_TRY(check_scan(c, k));
    const int P = c->P;
    std::vector<double> u0(P), pick(P);
    for (int p = 0; p < P; p++) {
        uint32_t w[4];
        Philox::draw(seed, (uint32_t)p, 0, kStreamClass, w);
        u0[p] = Philox::u01(w[0], w[1]) / nCls;
        pick[p] = Philox::u01(w[2], w[3]);
    }
    THX_TRY(draw_classes(c, nCls, u0.data(), pick.data(), clsOut));
    return from_global_impl(c, -1, clsOut, nCls, mLR, mLT, minStdR, minStdT, pfS, transS, seed);
}

int thx_pf_from_global_given(thx_ctx* c, int nCls, int mLR, int mLT, double minStdR,
                             double minStdT, double pfS, const double* transS,
                             const double* u0C, const double* pickC, const double* u0R,
                             const double* u0T, int* clsOut) {
    clear_error();
    THX_CHECK(c && transS && u0R && u0T && pfS > 0 && nCls >= 1 &&
                  (nCls == 1 || (u0C && pickC && clsOut)),
              THX_ERR_ARG, "pf_from_global_given: bad argument");
    for (int k = 0; k < nCls; k++) THX_TRY(check_scan(c, k));
    if (nCls == 1) {
        if (clsOut) std::fill(clsOut, clsOut + c->P, 0);
        return from_global_impl(c, 0, nullptr, 0, mLR, mLT, minStdR, minStdT, pfS, transS, 0,
                                u0R, u0T);
    }
    THX_TRY(draw_classes(c, nCls, u0C, pickC, clsOut));
    return from_global_impl(c, -1, clsOut, nCls, mLR, mLT, minStdR, minStdT, pfS, transS, 0,
                            u0R, u0T);
}

// Each active particle's n rotation samples and their weights, reordered in
// place along a Morton curve over (x, y, z) of the quaternion (w >= 0): the
// samples are an unordered set for the filter, and in this order the local
// scorer's 128-rotation work items are compact clusters whose gathers share
// cache lines.  One block per particle; bitonic sort of (key, index) in LDS;
// sets larger than kSortMaxN are left as they are.
constexpr int kSortMaxN = 1024;

__global__ __launch_bounds__(256) void k_sort_samples(double* __restrict__ q,
                                                      double* __restrict__ w,
                                                      float* __restrict__ wf, int n,
                                                      const int* __restrict__ active) {
    const int p = blockIdx.x, tid = threadIdx.x;
    if (n > kSortMaxN || (active && !active[p])) return;
    __shared__ uint64_t key[kSortMaxN];
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    double* qp = q + (size_t)p * n * 4;
    double* wp = w + (size_t)p * n;
    float* fp = wf + (size_t)p * n;
    for (int i = tid; i < n2; i += 256)
        key[i] = i < n ? ((uint64_t)quat_morton_dev(qp + 4 * i) << 32) | (uint32_t)i : ~0ull;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n2; i += 256) {
                const int x = i ^ j;
                if (x > i) {
                    const uint64_t a = key[i], b = key[x];
                    if ((a > b) == ((i & k) == 0)) {
                        key[i] = b;
                        key[x] = a;
                    }
                }
            }
            __syncthreads();
        }
    constexpr int S = kSortMaxN / 256;
    double qv[S][4], wv[S];
    float fv[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int i = tid + 256 * s;
        if (i < n) {
            const int src = (int)(key[i] & 0xffffffffu);
            for (int e = 0; e < 4; e++) qv[s][e] = qp[4 * src + e];
            wv[s] = wp[src];
            fv[s] = fp[src];
        }
    }
    __syncthreads();   // every source read before any write
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int i = tid + 256 * s;
        if (i < n) {
            for (int e = 0; e < 4; e++) qp[4 * i + e] = qv[s][e];
            wp[i] = wv[s];
            fp[i] = fv[s];
        }
    }
}

// Work list of a local-search phase, built on the device from the active
// flags (one block): the active particles in `order` (pose order), nbp work
// items of up to 128 rotations each, every particle's first item (bStart),
// its sample counts nRa = mLR / nTa = mLT (0 once converged), counts[1] =
// items, counts[2] = active particles.
__global__ __launch_bounds__(1024) void k_phase_items(const int* __restrict__ order,
                                                      const int* __restrict__ active, int P,
                                                      int mLR, int mLT, int nbp,
                                                      int2* __restrict__ blocks,
                                                      int* __restrict__ bStart,
                                                      int* __restrict__ nRa,
                                                      int* __restrict__ nTa,
                                                      int* __restrict__ counts) {
    __shared__ int part[1024];
    const int t = threadIdx.x, per = (P + 1023) / 1024;
    const int a = min(P, t * per), e = min(P, a + per);
    int n = 0;
    for (int i = a; i < e; i++) n += active[order[i]] != 0;
    part[t] = n;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {   // inclusive scan of the segment counts
        const int v = t >= s ? part[t - s] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int rank = part[t] - n;   // active particles in the segments before this one
    for (int i = a; i < e; i++) {
        const int p = order[i];
        const bool on = active[p] != 0;
        nRa[p] = on ? mLR : 0;
        nTa[p] = on ? mLT : 0;
        bStart[p] = rank * nbp;
        if (on) {
            for (int j = 0; j < nbp; j++)
                blocks[(size_t)rank * nbp + j] = make_int2(p, j * kLocalRotBlock);
            rank++;
        }
    }
    if (t == 1023) {
        bStart[P] = part[1023] * nbp;
        counts[1] = part[1023] * nbp;
        counts[2] = part[1023];
    }
}

// End of a phase (one block): counts[3] = particles still active, and the
// phase counter the next phase's kernels read advances.
__global__ __launch_bounds__(1024) void k_phase_end(const int* __restrict__ active, int P,
                                                    int* __restrict__ counts) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    int n = 0;
    for (int i = t; i < P; i += 1024) n += active[i] != 0;
    part[t] = n;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) part[t] += part[t + s];
        __syncthreads();
    }
    if (t == 0) {
        counts[3] = part[0];
        counts[0] += 1;
    }
}

__global__ void k_set_phase(int* __restrict__ counts, int phase) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[0] = phase;
}

namespace {
// The fixed inputs of one local search and the sample buffers of a phase
// (the current sets q / tr and the resampled ones q2 / tr2, which swap roles
// every phase).
struct RefineSetup {
    int P, mLR, mLT, nbp, TPAD, precision, mode, minPhase;
    uint64_t seed;
    double pfL, pfS;
    const int* pcls;
};
struct PhaseBufs {
    double *q, *q2, *tr, *tr2;
};
}  // namespace

// One local-search phase over every particle (call stack of updateSP,
// src/Optimiser.cpp:1139-1437, and the convergence test of :1475-1660),
// enqueued on the context stream with no host data: the work list comes from
// the active flags, the perturbation scale and random-stream phases from the
// device phase counter, so the same launches serve every phase (eagerly or
// as a replayed graph).
static int enqueue_phase(thx_ctx* c, const RefineSetup& r, const PhaseBufs& pb) {
    hipStream_t st = c->stream;
    const int P = r.P, mLR = r.mLR, mLT = r.mLT;
    int* dPh = c->pfCounts.as<int>();
    const int* dAct = c->pfActive.as<const int>();
    hipLaunchKernelGGL(k_phase_items, dim3(1), dim3(1024), 0, st, c->pfOrder.as<const int>(), dAct,
                       P, mLR, mLT, r.nbp, c->locBlocks.as<int2>(), c->locBStart.as<int>(),
                       c->pfNRa.as<int>(), c->pfNTa.as<int>(), dPh);
    // rotations: mean of the current samples, perturbation, balance + calVari
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, pb.q, mLR, dAct, 2, kAcgTol,
                       kAcgMaxIter, nullptr, c->pfMean.as<double>(), nullptr, nullptr, nullptr,
                       nullptr);
    THX_TRY(pf_perturb_rot(st, pb.q, c->pfOffR.as<int>(), c->pfNRa.as<int>(), mLR,
                           c->pfMean.as<double>(), c->pfK.as<double>(), 1.0,
                           scale_dev(dPh, r.pfL, r.pfS), P, r.seed, phase_dev(dPh, 2, 0)));
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, pb.q, mLR, dAct, 1 | 4 | 8, kAcgTol,
                       kAcgMaxIter, nullptr, nullptr, c->pfK.as<double>(), c->pfWR.as<double>(),
                       c->pfWRf.as<float>(), nullptr);
    hipLaunchKernelGGL(k_sort_samples, dim3(P), dim3(256), 0, st, pb.q, c->pfWR.as<double>(),
                       c->pfWRf.as<float>(), mLR, dAct);
    // shifts: perturbation, re-centring, balance + calVari
    hipLaunchKernelGGL(k_trans_stats, dim3(nb(P)), dim3(256), 0, st, pb.tr, mLT, dAct, P, 1,
                       c->pfS.as<double>(), scale_dev(dPh, r.pfL, r.pfS),
                       c->pfTransS.as<const double>(), kTransQ, r.seed, phase_dev(dPh, 2, 1),
                       nullptr, c->pfWT.as<double>(), c->pfWTf.as<float>());
    THX_HIP(hipGetLastError());
    // scoring of every active particle's samples (projection fused)
    LocalBatch b;
    b.vols = c->locVols.as<const float2* const>();
    b.volsLo = r.precision != THX_FP32 ? c->locVolsLo.as<const uint32_t* const>() : nullptr;
    b.pcls = r.pcls;
    b.quat = pb.q;
    b.pRf = c->pfWRf.as<const float>();
    b.rOff = c->pfOffR.as<const int>();
    b.nR = c->pfNRa.as<const int>();
    b.trans = pb.tr;
    b.pTf = c->pfWTf.as<const float>();
    b.tOff = c->pfOffT.as<const int>();
    b.nT = c->pfNTa.as<const int>();
    b.blocks = c->locBlocks.as<const int2>();
    b.bStart = c->locBStart.as<const int>();
    b.bMax = c->locBMax.as<float>();
    b.bWC = c->locBWC.as<float>();
    b.bArg = c->locBArg.as<int>();
    b.bWT = c->locBWT.as<float>();
    b.wRrel = c->locWRrel.as<float>();
    THX_TRY(launch_score_local(c, r.precision, r.TPAD, P * r.nbp, b, r.mode, dPh + 1));
    THX_TRY(launch_merge_local(c, P, kLocalWtStride, b, c->locWC.as<double>(),
                               c->locWR.as<double>(), c->locWT.as<double>(),
                               c->locBase.as<double>(), c->locBestR.as<int>(),
                               c->locBestT.as<int>()));
    // half-height peak of the rotation marginal, top samples
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, c->locWR.as<double>(), mLR, dAct,
                       c->pfPeak.as<double>(), 0, 1, (const double*)pb.q, 4,
                       c->pfTopR.as<double>(), nullptr);
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, c->locWT.as<double>(), mLT, dAct,
                       c->pfPeak.as<double>(), 0, 0, (const double*)pb.tr, 2,
                       c->pfTopT.as<double>(), nullptr);
    // resampling of both sets into q2 / tr2
    THX_TRY(pf_resample(st, c->pfWR.as<double>(), c->locWR.as<double>(), c->pfOffR.as<int>(),
                        c->pfNRa.as<int>(), c->pfOffR.as<int>(), c->pfNRa.as<int>(), P, r.seed,
                        phase_dev(dPh, 2, 1000), c->pfIdx.as<int>(), c->pfWNew.as<double>()));
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLR)), dim3(256), 0, st,
                       (const double*)pb.q, mLR, 0, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLR, 4, dAct, P, pb.q2, c->pfWR.as<double>(),
                       c->pfWRf.as<float>());
    THX_TRY(pf_resample(st, c->pfWT.as<double>(), c->locWT.as<double>(), c->pfOffT.as<int>(),
                        c->pfNTa.as<int>(), c->pfOffT.as<int>(), c->pfNTa.as<int>(), P, r.seed,
                        phase_dev(dPh, 2, 1001), c->pfIdx.as<int>(), c->pfWNew.as<double>()));
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLT)), dim3(256), 0, st,
                       (const double*)pb.tr, mLT, 0, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLT, 2, dAct, P, pb.tr2, c->pfWT.as<double>(),
                       c->pfWTf.as<float>());
    // converged particles were gathered neither way: their rows are copied
    hipLaunchKernelGGL(k_keep_rows, dim3(nb((size_t)P * mLR * 4)), dim3(256), 0, st,
                       (const double*)pb.q, pb.q2, mLR * 4, dAct, P);
    hipLaunchKernelGGL(k_keep_rows, dim3(nb((size_t)P * mLT * 2)), dim3(256), 0, st,
                       (const double*)pb.tr, pb.tr2, mLT * 2, dAct, P);
    hipLaunchKernelGGL(k_converge, dim3(nb(P)), dim3(256), 0, st, c->pfK.as<const double>(),
                       c->pfS.as<const double>(), c->pfActive.as<int>(), c->pfBest.as<double>(),
                       c->pfNPhase.as<int>(), P, phase_dev(dPh, 1, 0), r.minPhase);
    hipLaunchKernelGGL(k_phase_end, dim3(1), dim3(1024), 0, st, dAct, P, dPh);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// The phase graph for this setup and buffer assignment: reused while every
// pointer and scalar it holds is unchanged, else captured (at most four are
// kept: the two buffer assignments of up to two setups).
static int phase_graph(thx_ctx* c, const RefineSetup& r, const PhaseBufs& pb,
                       hipGraphExec_t* out) {
    std::vector<long long> key;
    auto add = [&](const void* x) { key.push_back((long long)(uintptr_t)x); };
    for (const void* x : std::initializer_list<const void*>{pb.q, pb.q2, pb.tr,
                          pb.tr2, c->pfWR.p, c->pfWT.p, c->pfWRf.p, c->pfWTf.p,
                          c->pfIdx.p, c->pfWNew.p, c->pfK.p, c->pfS.p, c->pfMean.p, c->pfTopR.p,
                          c->pfTopT.p, c->pfPeak.p, c->pfTransS.p, c->pfBest.p, c->pfActive.p,
                          c->pfNPhase.p, c->pfOffR.p, c->pfOffT.p, c->pfNRa.p, c->pfNTa.p,
                          c->pfOrder.p, c->pfCounts.p, c->locBlocks.p, c->locBStart.p,
                          c->locBMax.p, c->locBWC.p, c->locBArg.p, c->locBWT.p, c->locWRrel.p,
                          c->locWC.p, c->locWR.p, c->locWT.p, c->locBase.p, c->locBestR.p,
                          c->locBestT.p, c->locVols.p, c->locVolsLo.p, r.pcls,
                          c->Y.p, c->Wt.p, c->Aterm.p, c->iCol.p, c->iRow.p, c->iSig.p,
                          c->shellScale.p, c->locCounter.p})
        add(x);
    double d[2] = {r.pfL, r.pfS};
    long long dd[2];
    std::memcpy(dd, d, sizeof dd);
    for (long long v : std::initializer_list<long long>{(long long)r.P, (long long)r.mLR, (long long)r.mLT, (long long)r.TPAD,
                        (long long)r.precision, (long long)r.mode, (long long)r.minPhase,
                        (long long)r.seed, dd[0], dd[1], (long long)c->K, (long long)c->Kp,
                        (long long)c->idim, (long long)c->pf})
        key.push_back(v);
    for (auto& g : c->pfGraphs)
        if (g.first == key) {
            *out = g.second;
            return THX_OK;
        }
    hipStream_t st = c->stream;
    THX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int cs = enqueue_phase(c, r, pb);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &g);
    if (cs != THX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return cs;
    }
    THX_CHECK(ee == hipSuccess && g, THX_ERR_HIP, "pf_refine: graph capture failed");
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    THX_CHECK(ie == hipSuccess, THX_ERR_HIP, "pf_refine: graph instantiation failed");
    if (c->pfGraphs.size() >= 4) {
        (void)hipGraphExecDestroy(c->pfGraphs.front().second);
        c->pfGraphs.erase(c->pfGraphs.begin());
    }
    c->pfGraphs.push_back({key, ge});
    *out = ge;
    return THX_OK;
}

int thx_pf_refine(thx_ctx* c, const int* cls, int precision, int afterGlobal, double pfL,
                  double pfS, int maxPhase, uint64_t seed, int* phasesRun) {
    clear_error();
    THX_CHECK(c && c->pfValid && c->pfP == c->P, THX_ERR_STATE,
              "pf_refine: particle-filter state not set for the context's particles");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16 || precision == THX_FP16,
              THX_ERR_ARG, "pf_refine: precision must be FP32, BF16 or FP16");
    THX_CHECK(!c->vol.empty() && c->K > 0, THX_ERR_STATE, "pf_refine: volume and pixels needed");
    const int nCls = (int)c->vol.size();
    // every class the particles use must hold a volume of this pixel set's size
    std::vector<char> used(nCls, 0);
    if (cls) {
        for (int q = 0; q < c->P; q++)
            if (cls[q] >= 0 && cls[q] < nCls) used[cls[q]] = 1;
    } else {
        used[0] = 1;
    }
    for (int k = 0; k < nCls; k++)
        THX_CHECK(!used[k] || c->vdim[k] == c->idim * c->pf, THX_ERR_STATE,
                  "pf_refine: a class volume in use does not match idim * pf");
    THX_HIP(hipSetDevice(c->device));
    const int P = c->P, mLR = c->pfR, mLT = c->pfT;
    if (cls)
        for (int p = 0; p < P; p++)
            THX_CHECK(cls[p] >= 0 && cls[p] < nCls, THX_ERR_ARG, "pf_refine: bad class");
    hipStream_t st = c->stream;
    // scorer inputs that stay fixed over the phases
    std::vector<const float2*> vp(nCls);
    for (int k = 0; k < nCls; k++) vp[k] = c->vol[k].as<const float2>();
    THX_TRY(upload(c->locVols, vp.data(), vp.size(), st));
    if (precision != THX_FP32) {
        THX_TRY(ensure_vol_lo(c, used));
        std::vector<const uint32_t*> vl(nCls, nullptr);
        for (int k = 0; k < nCls; k++)
            if (used[k]) vl[k] = c->volLo[k].as<const uint32_t>();
        THX_TRY(upload(c->locVolsLo, vl.data(), vl.size(), st));
    }
    if (cls) THX_TRY(upload(c->locPcls, cls, P, st));
    RefineSetup r;
    r.P = P;
    r.mLR = mLR;
    r.mLT = mLT;
    r.TPAD = kLocalTile[local_tile_index(mLT)];
    r.nbp = (mLR + kLocalRotBlock - 1) / kLocalRotBlock;
    r.precision = precision;
    r.minPhase = afterGlobal ? 10 : 3;   // MIN_N_PHASE_PER_ITER_{GLOBAL,LOCAL}
    r.seed = seed;
    r.pfL = pfL;
    r.pfS = pfS;
    r.pcls = cls ? c->locPcls.as<const int>() : nullptr;
    const size_t nItemMax = (size_t)P * r.nbp;
    THX_CHECK(nItemMax < (1u << 31), THX_ERR_RANGE, "pf_refine: too many work items");
    THX_TRY(c->locBMax.ensure(sizeof(float) * nItemMax));
    THX_TRY(c->locBWC.ensure(sizeof(float) * nItemMax));
    THX_TRY(c->locBArg.ensure(sizeof(int) * nItemMax));
    THX_TRY(c->locBWT.ensure(sizeof(float) * nItemMax * kLocalWtStride));
    THX_TRY(c->locBlocks.ensure(sizeof(int2) * nItemMax));
    THX_TRY(c->locBStart.ensure(sizeof(int) * (P + 1)));
    THX_TRY(c->locWRrel.ensure(sizeof(float) * (size_t)P * mLR));
    THX_TRY(c->locWC.ensure(sizeof(double) * P));
    THX_TRY(c->locWR.ensure(sizeof(double) * (size_t)P * mLR));
    THX_TRY(c->locWT.ensure(sizeof(double) * (size_t)P * mLT));
    THX_TRY(c->locBase.ensure(sizeof(double) * P));
    THX_TRY(c->locBestR.ensure(sizeof(int) * P));
    THX_TRY(c->locBestT.ensure(sizeof(int) * P));
    THX_TRY(c->pfNRa.ensure(sizeof(int) * P));
    THX_TRY(c->pfNTa.ensure(sizeof(int) * P));
    THX_TRY(c->pfCounts.ensure(4 * sizeof(int)));
    THX_TRY(c->pfCountsH.ensure(4 * sizeof(int)));
    THX_TRY(c->locCounter.ensure(8 * sizeof(int)));   // no allocation inside a capture
    c->locValid = false;   // the loc* buffers now hold this search's phases
    // FP16: per-shell operand scales (particles and volumes stay fixed over the phases)
    if (precision == THX_FP16) THX_TRY(launch_local_shell_scale(c, used));
    THX_TRY(launch_prep_particles(c, precision == THX_FP16 ? THX_FP16 : THX_FP32));
    const int start = afterGlobal ? 1 : 0;
    // work-list order by pose (the particles move little over the phases):
    // the first sample of each particle
    std::vector<double> q0(4 * (size_t)P);
    std::vector<int> off0(P);
    THX_HIP(hipMemcpy2D(q0.data(), 4 * sizeof(double), c->pfQ.p, 4 * sizeof(double) * mLR,
                        4 * sizeof(double), P, hipMemcpyDeviceToHost));
    for (int p = 0; p < P; p++) off0[p] = p;
    const std::vector<int> order = pose_order(q0.data(), off0.data(), cls, P);
    THX_TRY(upload(c->pfOrder, order.data(), P, st));
    // the first phase's work list, for the launch-shape tuner
    int* cH = c->pfCountsH.as<int>();
    int* dPh = c->pfCounts.as<int>();
    hipLaunchKernelGGL(k_set_phase, dim3(1), dim3(64), 0, st, dPh, start);
    hipLaunchKernelGGL(k_phase_items, dim3(1), dim3(1024), 0, st, c->pfOrder.as<const int>(),
                       c->pfActive.as<const int>(), P, mLR, mLT, r.nbp, c->locBlocks.as<int2>(),
                       c->locBStart.as<int>(), c->pfNRa.as<int>(), c->pfNTa.as<int>(), dPh);
    THX_HIP(hipGetLastError());
    THX_HIP(hipMemcpyAsync(cH, dPh, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    THX_HIP(hipStreamSynchronize(st));
    int phase = start, ran = 0;
    if (cH[2] > 0 && phase < maxPhase) {
        // the first phase's current samples, scored by both launch shapes
        LocalBatch b;
        b.vols = c->locVols.as<const float2* const>();
        b.volsLo = precision != THX_FP32 ? c->locVolsLo.as<const uint32_t* const>() : nullptr;
        b.pcls = r.pcls;
        b.quat = c->pfQ.as<const double>();
        b.pRf = c->pfWRf.as<const float>();
        b.rOff = c->pfOffR.as<const int>();
        b.nR = c->pfNRa.as<const int>();
        b.trans = c->pfTr.as<const double>();
        b.pTf = c->pfWTf.as<const float>();
        b.tOff = c->pfOffT.as<const int>();
        b.nT = c->pfNTa.as<const int>();
        b.blocks = c->locBlocks.as<const int2>();
        b.bStart = c->locBStart.as<const int>();
        b.bMax = c->locBMax.as<float>();
        b.bWC = c->locBWC.as<float>();
        b.bArg = c->locBArg.as<int>();
        b.bWT = c->locBWT.as<float>();
        b.wRrel = c->locWRrel.as<float>();
        r.mode = choose_local_mode(c, precision, r.TPAD, cH[1], b);
        if (r.mode < 0) return THX_ERR_HIP;
        c->locModeUsed = r.mode;
    }
    // the phases: the same launches every time, replayed as a graph when
    // graphs are on; the sample buffers swap roles every phase
    for (; phase < maxPhase && cH[2] > 0; phase++) {
        const bool odd = ((phase - start) & 1) != 0;
        PhaseBufs pb;
        pb.q = odd ? c->pfQ2.as<double>() : c->pfQ.as<double>();
        pb.q2 = odd ? c->pfQ.as<double>() : c->pfQ2.as<double>();
        pb.tr = odd ? c->pfTr2.as<double>() : c->pfTr.as<double>();
        pb.tr2 = odd ? c->pfTr.as<double>() : c->pfTr2.as<double>();
        if (c->locGraphOn) {
            hipGraphExec_t g = nullptr;
            THX_TRY(phase_graph(c, r, pb, &g));
            THX_HIP(hipGraphLaunch(g, st));
        } else {
            THX_TRY(enqueue_phase(c, r, pb));
        }
        THX_HIP(hipMemcpyAsync(cH, dPh, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
        THX_HIP(hipStreamSynchronize(st));
        ran++;
        cH[2] = cH[3];   // the next phase's active particles
    }
    // the current sample sets are in pfQ / pfTr again
    if (ran & 1) {
        std::swap(c->pfQ, c->pfQ2);
        std::swap(c->pfTr, c->pfTr2);
    }
    if (phasesRun) *phasesRun = ran;
    c->pfLastPhase = phase;
    return THX_OK;
}

// Effective sample size of each particle's row of n weights, (sum w)^2 /
// sum w^2 (the commented-out Particle::neff, src/Particle.cpp:1647-1651, of
// the normalised weights).  One block per particle.
__global__ __launch_bounds__(256) void k_row_ess(const double* __restrict__ w, int n, int P,
                                                 double* __restrict__ ess) {
    const int p = blockIdx.x;
    if (p >= P) return;
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double v = w[(size_t)p * n + i];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ess[p] = s2[0] > 0 ? s1[0] * s1[0] / s2[0] : 0.0;
}

}  // extern "C"

// dst row idx[i] = src row i (rows of rowLen elements), i < n
template <class T>
__global__ void k_scatter_rows(const T* __restrict__ src, int rowLen, const int* __restrict__ idx,
                               int n, T* __restrict__ dst) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * rowLen) return;
    const size_t i = e / rowLen, j = e % rowLen;
    dst[(size_t)idx[i] * rowLen + j] = src[e];
}

template <class T>
static int scatter_rows(hipStream_t st, const DevBuf& src, int rowLen, const int* idx, int n,
                        DevBuf& dst) {
    hipLaunchKernelGGL(k_scatter_rows<T>, dim3(nb((size_t)n * rowLen)), dim3(256), 0, st,
                       src.as<const T>(), rowLen, idx, n, dst.as<T>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

extern "C" {

// Particle-filter states of the particles idx from their whole-grid scan of
// class cls (the marginals scan_subset left in lpRes): the post-scan update
// of thx_pf_from_global computed for those rows alone, then scattered into
// the context's full state arrays.
static int states_from_subset_scan(thx_ctx* c, int cls, const std::vector<int>& idx, int mLR,
                                   int mLT, double minStdR, double minStdT, double pfS,
                                   const std::vector<double>& transSAll, uint64_t seed) {
    const int n = (int)idx.size(), P0 = c->P, Ppad0 = c->Ppad;
    hipStream_t st = c->stream;
    DevBuf dIdx;
    THX_TRY(upload(dIdx, idx.data(), n, st));
    std::vector<double> tS(n);
    for (int i = 0; i < n; i++) tS[i] = transSAll[idx[i]];
    // the full state arrays step aside; from_global_impl fills fresh ones for
    // the n rows (scratch buffers are shared)
    DevBuf* state[] = {&c->pfQ, &c->pfTr, &c->pfWR, &c->pfWRf, &c->pfWT, &c->pfWTf, &c->pfK,
                       &c->pfS, &c->pfPeak, &c->pfMean, &c->pfTransS};
    constexpr int NS = sizeof(state) / sizeof(state[0]);
    DevBuf saved[NS];
    for (int i = 0; i < NS; i++) std::swap(saved[i], *state[i]);
    std::swap(c->res[cls], c->lpRes);
    c->P = n;
    c->Ppad = round_up(n, 128);
    int s = from_global_impl(c, cls, nullptr, 1, mLR, mLT, minStdR, minStdT, pfS, tS.data(), seed);
    c->P = P0;
    c->Ppad = Ppad0;
    std::swap(c->res[cls], c->lpRes);
    if (s == THX_OK) {
        const int* di = dIdx.as<const int>();
        const int lens[NS] = {4 * mLR, 2 * mLT, mLR, mLR, mLT, mLT, 3, 3, 1, 4, 1};
        for (int i = 0; i < NS && s == THX_OK; i++) {
            if (state[i] == &c->pfWRf || state[i] == &c->pfWTf)
                s = scatter_rows<float>(st, *state[i], lens[i], di, n, saved[i]);
            else
                s = scatter_rows<double>(st, *state[i], lens[i], di, n, saved[i]);
        }
        if (s == THX_OK && hipStreamSynchronize(st) != hipSuccess) {
            set_error("seeded: scattering the fallback states failed");
            s = THX_ERR_HIP;
        }
    }
    for (int i = 0; i < NS; i++) {
        state[i]->release();
        std::swap(saved[i], *state[i]);
    }
    dIdx.release();
    c->pfP = P0;
    c->pfR = mLR;
    c->pfT = mLT;
    return s;
}

int thx_ctx_expect_seeded(thx_ctx* c, thx_cnn* net, int cls, int mLR, int mLT, int precision,
                          double essMin, int fallbackPrecision, double minStdR, double minStdT,
                          double pfL, double pfS, const double* transS, int maxPhase,
                          uint64_t seed, int* nFallback, int* phasesRun) {
    clear_error();
    THX_CHECK(c && net && transS, THX_ERR_ARG, "expect_seeded: null argument");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16 || precision == THX_FP16,
              THX_ERR_ARG, "expect_seeded: precision must be FP32, BF16 or FP16");
    THX_CHECK(mLR >= 2 && mLR <= kResampleMaxN && mLT >= 2 && mLT <= 128, THX_ERR_RANGE,
              "expect_seeded: sample counts out of range");
    THX_CHECK(c->K > 0 && c->P > 0 && cls >= 0 && cls < (int)c->vol.size() &&
                  c->vdim[cls] == c->idim * c->pf,
              THX_ERR_STATE, "expect_seeded: pixels, particles and the class volume needed");
    THX_CHECK(c->ctfReady && c->sigReady, THX_ERR_STATE,
              "expect_seeded: particle CTF and noise spectrum must be set");
    int C = 0;
    THX_TRY(thx_cnn_components(net, &C));
    THX_HIP(hipSetDevice(c->device));
    const int P = c->P;
    hipStream_t st = c->stream;
    std::vector<double> tS(transS, transS + P);
    // 1. the proposal's mixture for every particle, then mLR rotations and mLT
    //    shifts drawn from it: the particle filter's samples (Particle::reset
    //    / load draw them from ACG distributions, src/Particle.cpp:87-169,
    //    401-554)
    THX_TRY(pf_alloc(c, P, mLR, mLT));
    THX_HIP(hipMemcpy(c->pfTransS.p, tS.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    THX_TRY(c->lpMix.ensure(sizeof(double) * (size_t)P * (6 * C + 3)));
    if (!c->lpReady) THX_HIP(hipEventCreateWithFlags(&c->lpReady, hipEventDisableTiming));
    if (!c->lpDone) THX_HIP(hipEventCreateWithFlags(&c->lpDone, hipEventDisableTiming));
    THX_HIP(hipEventRecord(c->lpReady, st));
    THX_TRY(cnn_propose_async(net, c, c->lpMix.as<double>(), c->lpReady, c->lpDone));
    THX_HIP(hipStreamWaitEvent(st, c->lpDone, 0));
    THX_TRY(pf_sample_mixture(st, c->lpMix.as<const double>(), C, P, mLR, mLT, seed, 0u,
                              c->pfQ.as<double>(), c->pfWNew.as<double>(), c->pfTr.as<double>()));
    // 2. the samples' statistics: ACG mean and concentrations, balance weights
    //    (the sample density's inverse), shift means and deviations
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, c->pfQ.as<const double>(), mLR,
                       nullptr, 1 | 2 | 4 | 8, kAcgTol, kAcgMaxIter, nullptr, c->pfMean.as<double>(),
                       c->pfK.as<double>(), c->pfWR.as<double>(), c->pfWRf.as<float>(), nullptr);
    hipLaunchKernelGGL(k_sort_samples, dim3(P), dim3(256), 0, st, c->pfQ.as<double>(),
                       c->pfWR.as<double>(), c->pfWRf.as<float>(), mLR, (const int*)nullptr);
    hipLaunchKernelGGL(k_trans_stats, dim3(nb(P)), dim3(256), 0, st, c->pfTr.as<double>(), mLT,
                       nullptr, P, 0, c->pfS.as<double>(), scale_host(0.0),
                       c->pfTransS.as<const double>(), kTransQ, seed, phase_host(0u), nullptr,
                       c->pfWT.as<double>(), c->pfWTf.as<float>());
    THX_HIP(hipGetLastError());
    // 3. one scoring pass of the seeded samples (the local scorer, every
    //    particle in one launch), the peak factor and top samples, the ESS of
    //    each particle's rotation weights
    std::vector<char> used(c->vol.size(), 0);
    used[cls] = 1;
    std::vector<const float2*> vp(c->vol.size());
    for (size_t k = 0; k < c->vol.size(); k++) vp[k] = c->vol[k].as<const float2>();
    THX_TRY(upload(c->locVols, vp.data(), vp.size(), st));
    if (precision != THX_FP32) {
        THX_TRY(ensure_vol_lo(c, used));
        std::vector<const uint32_t*> vl(c->vol.size(), nullptr);
        vl[cls] = c->volLo[cls].as<const uint32_t>();
        THX_TRY(upload(c->locVolsLo, vl.data(), vl.size(), st));
    }
    std::vector<int> clsv(P, cls);
    THX_TRY(upload(c->locPcls, clsv.data(), P, st));
    if (precision == THX_FP16) THX_TRY(launch_local_shell_scale(c, used));
    THX_TRY(launch_prep_particles(c, precision == THX_FP16 ? THX_FP16 : THX_FP32));
    const int nbp = (mLR + kLocalRotBlock - 1) / kLocalRotBlock;
    const size_t nItem = (size_t)P * nbp;
    THX_TRY(c->locBMax.ensure(sizeof(float) * nItem));
    THX_TRY(c->locBWC.ensure(sizeof(float) * nItem));
    THX_TRY(c->locBArg.ensure(sizeof(int) * nItem));
    THX_TRY(c->locBWT.ensure(sizeof(float) * nItem * kLocalWtStride));
    THX_TRY(c->locWRrel.ensure(sizeof(float) * (size_t)P * mLR));
    THX_TRY(c->locWC.ensure(sizeof(double) * P));
    THX_TRY(c->locWR.ensure(sizeof(double) * (size_t)P * mLR));
    THX_TRY(c->locWT.ensure(sizeof(double) * (size_t)P * mLT));
    THX_TRY(c->locBase.ensure(sizeof(double) * P));
    THX_TRY(c->locBestR.ensure(sizeof(int) * P));
    THX_TRY(c->locBestT.ensure(sizeof(int) * P));
    std::vector<int2> blocks;
    std::vector<int> bStart(P + 1);
    for (int p = 0; p < P; p++) {
        bStart[p] = (int)blocks.size();
        for (int r0 = 0; r0 < mLR; r0 += kLocalRotBlock) blocks.push_back(make_int2(p, r0));
    }
    bStart[P] = (int)blocks.size();
    THX_TRY(upload(c->locBlocks, blocks.data(), blocks.size(), st));
    THX_TRY(upload(c->locBStart, bStart.data(), P + 1, st));
    LocalBatch b;
    b.vols = c->locVols.as<const float2* const>();
    b.volsLo = precision != THX_FP32 ? c->locVolsLo.as<const uint32_t* const>() : nullptr;
    b.pcls = c->locPcls.as<const int>();
    b.quat = c->pfQ.as<const double>();
    b.pRf = c->pfWRf.as<const float>();
    b.rOff = c->pfOffR.as<const int>();
    b.nR = c->pfNR.as<const int>();
    b.trans = c->pfTr.as<const double>();
    b.pTf = c->pfWTf.as<const float>();
    b.tOff = c->pfOffT.as<const int>();
    b.nT = c->pfNT.as<const int>();
    b.blocks = c->locBlocks.as<const int2>();
    b.bStart = c->locBStart.as<const int>();
    b.bMax = c->locBMax.as<float>();
    b.bWC = c->locBWC.as<float>();
    b.bArg = c->locBArg.as<int>();
    b.bWT = c->locBWT.as<float>();
    b.wRrel = c->locWRrel.as<float>();
    const int TPAD = kLocalTile[local_tile_index(mLT)];
    const int mode = choose_local_mode(c, precision, TPAD, (int)blocks.size(), b);
    if (mode < 0) return THX_ERR_HIP;
    THX_TRY(launch_score_local(c, precision, TPAD, (int)blocks.size(), b, mode));
    THX_TRY(launch_merge_local(c, P, kLocalWtStride, b, c->locWC.as<double>(),
                               c->locWR.as<double>(), c->locWT.as<double>(),
                               c->locBase.as<double>(), c->locBestR.as<int>(),
                               c->locBestT.as<int>()));
    DevBuf dEss;
    THX_TRY(dEss.ensure(sizeof(double) * P));
    hipLaunchKernelGGL(k_row_ess, dim3(P), dim3(256), 0, st, c->locWR.as<const double>(), mLR, P,
                       dEss.as<double>());
    // setPeakFactor + keepHalfHeightPeak of the scored rotations, tops
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, c->locWR.as<double>(), mLR,
                       nullptr, c->pfPeak.as<double>(), 1, 1, c->pfQ.as<const double>(), 4,
                       c->pfTopR.as<double>(), nullptr);
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, c->locWT.as<double>(), mLT,
                       nullptr, c->pfPeak.as<double>(), 0, 0, c->pfTr.as<const double>(), 2,
                       c->pfTopT.as<double>(), nullptr);
    // 4. resampling of both sets by their scores, calVari, scanning minima:
    //    the particle-filter state the local phases start from
    THX_TRY(pf_resample(st, c->pfWR.as<double>(), c->locWR.as<double>(), c->pfOffR.as<int>(),
                        c->pfNR.as<int>(), c->pfOffR.as<int>(), c->pfNR.as<int>(), P, seed,
                        phase_host(1u), c->pfIdx.as<int>(), c->pfWNew.as<double>()));
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLR)), dim3(256), 0, st,
                       c->pfQ.as<const double>(), mLR, 0, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLR, 4, nullptr, P, c->pfQ2.as<double>(),
                       c->pfWR.as<double>(), c->pfWRf.as<float>());
    THX_TRY(pf_resample(st, c->pfWT.as<double>(), c->locWT.as<double>(), c->pfOffT.as<int>(),
                        c->pfNT.as<int>(), c->pfOffT.as<int>(), c->pfNT.as<int>(), P, seed,
                        phase_host(2u), c->pfIdx.as<int>(), c->pfWNew.as<double>()));
    hipLaunchKernelGGL(k_gather_samples, dim3(nb((size_t)P * mLT)), dim3(256), 0, st,
                       c->pfTr.as<const double>(), mLT, 0, c->pfIdx.as<const int>(),
                       c->pfWNew.as<const double>(), mLT, 2, nullptr, P, c->pfTr2.as<double>(),
                       c->pfWT.as<double>(), c->pfWTf.as<float>());
    std::swap(c->pfQ, c->pfQ2);
    std::swap(c->pfTr, c->pfTr2);
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, c->pfQ.as<const double>(), mLR,
                       nullptr, 6, kAcgTol, kAcgMaxIter, nullptr, c->pfMean.as<double>(),
                       c->pfK.as<double>(), nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_trans_stats, dim3(nb(P)), dim3(256), 0, st, c->pfTr.as<double>(), mLT,
                       nullptr, P, 0, c->pfS.as<double>(), scale_host(0.0),
                       c->pfTransS.as<const double>(), kTransQ, seed, phase_host(0u), nullptr,
                       nullptr, nullptr);
    const double kMin = (minStdR / pfS) * (minStdR / pfS), sMin = minStdT / pfS;
    hipLaunchKernelGGL(k_min_std, dim3(nb(P)), dim3(256), 0, st, c->pfK.as<double>(),
                       c->pfS.as<double>(), P, kMin, sMin);
    THX_HIP(hipGetLastError());
    std::vector<double> ess(P);
    THX_HIP(hipMemcpyAsync(ess.data(), dEss.p, sizeof(double) * P, hipMemcpyDeviceToHost, st));
    THX_HIP(hipStreamSynchronize(st));
    dEss.release();
    // 5. the uniform global search where the weights collapsed: whole-grid
    //    scan of those particles and their states from it
    std::vector<int> fell;
    if (essMin > 0)
        for (int p = 0; p < P; p++)
            if (!(ess[p] >= essMin)) fell.push_back(p);
    if (!fell.empty()) {
        THX_CHECK(c->R > 0 && c->T > 0 && c->rotP.p && c->normP.p &&
                      c->rotP.bytes >= sizeof(float2) * (size_t)c->Rpad * c->Kp &&
                      c->normP.bytes >= sizeof(float) * (size_t)c->Rpad * c->Kp,
                  THX_ERR_STATE, "expect_seeded: the fallback needs the context's grid");
        // the grid's slices of this class (the scan reads the projections)
        THX_TRY(launch_project(c, cls, c->quat.as<const double>(), c->R, c->Rpad,
                               c->rotP.as<float2>(), c->normP.as<float>()));
        THX_TRY(scan_subset(c, cls, fallbackPrecision, fell));
        THX_TRY(states_from_subset_scan(c, cls, fell, mLR, mLT, minStdR, minStdT, pfS, tS,
                                        seed ^ 0x5eedull));
    }
    if (nFallback) *nFallback = (int)fell.size();
    c->pfValid = true;
    // 6. the local phases for every particle
    return thx_pf_refine(c, cls ? clsv.data() : nullptr, precision, 1, pfL, pfS, maxPhase,
                         seed + 1, phasesRun);
}

int thx_pf_get_state(thx_ctx* c, double* quat, double* trans, double* wR, double* wT, double* k,
                     double* s01, double* topR, double* topT, double* peakR, int* nPhase) {
    clear_error();
    THX_CHECK(c && c->pfValid, THX_ERR_STATE, "pf_get_state: no particle-filter state");
    THX_HIP(hipSetDevice(c->device));
    THX_HIP(hipStreamSynchronize(c->stream));
    const size_t P = c->pfP, nr = P * c->pfR, nt = P * c->pfT;
    if (quat) THX_HIP(hipMemcpy(quat, c->pfQ.p, sizeof(double) * 4 * nr, hipMemcpyDeviceToHost));
    if (trans) THX_HIP(hipMemcpy(trans, c->pfTr.p, sizeof(double) * 2 * nt, hipMemcpyDeviceToHost));
    if (wR) THX_HIP(hipMemcpy(wR, c->pfWR.p, sizeof(double) * nr, hipMemcpyDeviceToHost));
    if (wT) THX_HIP(hipMemcpy(wT, c->pfWT.p, sizeof(double) * nt, hipMemcpyDeviceToHost));
    if (k) THX_HIP(hipMemcpy(k, c->pfK.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    if (s01) {
        std::vector<double> s3(3 * P);
        THX_HIP(hipMemcpy(s3.data(), c->pfS.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) {
            s01[2 * p] = s3[3 * p];
            s01[2 * p + 1] = s3[3 * p + 1];
        }
    }
    if (topR) THX_HIP(hipMemcpy(topR, c->pfTopR.p, sizeof(double) * 4 * P, hipMemcpyDeviceToHost));
    if (topT) THX_HIP(hipMemcpy(topT, c->pfTopT.p, sizeof(double) * 2 * P, hipMemcpyDeviceToHost));
    if (peakR) THX_HIP(hipMemcpy(peakR, c->pfPeak.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    if (nPhase) THX_HIP(hipMemcpy(nPhase, c->pfNPhase.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    return THX_OK;
}

// Stage kernels on host buffers, for checking: rotation statistics of P sets
// of n samples (A[16], mean[4], k[3], balance weights[n], iterations).
int thx_pf_rot_stats(thx_ctx* c, const double* quat, int n, int P, double* A, double* mean,
                     double* k, double* wBal, int* iters) {
    clear_error();
    THX_CHECK(c && quat && n >= 1 && P >= 1, THX_ERR_ARG, "pf_rot_stats: bad argument");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dq, dA, dm, dk, dw, di;
    const size_t N = (size_t)P * n;
    int s = THX_OK;
    if ((s = upload(dq, quat, 4 * N, st)) || (s = dA.ensure(sizeof(double) * 16 * P)) ||
        (s = dm.ensure(sizeof(double) * 4 * P)) || (s = dk.ensure(sizeof(double) * 3 * P)) ||
        (s = dw.ensure(sizeof(double) * N)) || (s = di.ensure(sizeof(int) * P)))
        return s;
    hipLaunchKernelGGL(k_rot_stats, dim3(P), dim3(256), 0, st, dq.as<const double>(), n, nullptr,
                       1 | 2 | 4, kAcgTol, kAcgMaxIter, dA.as<double>(), dm.as<double>(),
                       dk.as<double>(), dw.as<double>(), nullptr, di.as<int>());
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    if (A) THX_HIP(hipMemcpy(A, dA.p, sizeof(double) * 16 * P, hipMemcpyDeviceToHost));
    if (mean) THX_HIP(hipMemcpy(mean, dm.p, sizeof(double) * 4 * P, hipMemcpyDeviceToHost));
    if (k) THX_HIP(hipMemcpy(k, dk.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    if (wBal) THX_HIP(hipMemcpy(wBal, dw.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (iters) THX_HIP(hipMemcpy(iters, di.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    DevBuf* all[] = {&dq, &dA, &dm, &dk, &dw, &di};
    for (DevBuf* b : all) b->release();
    return THX_OK;
}

// Shift statistics of P sets of n samples: means[2], deviations[2] (n - 1)
// and balance weights[n].
int thx_pf_trans_stats(thx_ctx* c, const double* trans, int n, int P, double* m, double* s01,
                       double* wBal) {
    clear_error();
    THX_CHECK(c && trans && n >= 2 && P >= 1, THX_ERR_ARG, "pf_trans_stats: bad argument");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dt, dm, ds, dw;
    const size_t N = (size_t)P * n;
    int s = THX_OK;
    if ((s = upload(dt, trans, 2 * N, st)) || (s = dm.ensure(sizeof(double) * 2 * P)) ||
        (s = ds.ensure(sizeof(double) * 3 * P)) || (s = dw.ensure(sizeof(double) * N)))
        return s;
    hipLaunchKernelGGL(k_trans_stats, dim3(nb(P)), dim3(256), 0, st, dt.as<double>(), n, nullptr,
                       P, 0, ds.as<double>(), scale_host(0.0), nullptr, kTransQ, 0ull,
                       phase_host(0u), dm.as<double>(), dw.as<double>(), nullptr);
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    std::vector<double> s3(3 * (size_t)P);
    THX_HIP(hipMemcpy(s3.data(), ds.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    if (s01)
        for (int p = 0; p < P; p++) {
            s01[2 * p] = s3[3 * p];
            s01[2 * p + 1] = s3[3 * p + 1];
        }
    if (m) THX_HIP(hipMemcpy(m, dm.p, sizeof(double) * 2 * P, hipMemcpyDeviceToHost));
    if (wBal) THX_HIP(hipMemcpy(wBal, dw.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    DevBuf* all[] = {&dt, &dm, &ds, &dw};
    for (DevBuf* b : all) b->release();
    return THX_OK;
}

// Convergence test over scripted per-phase statistics (checking): phase i
// (phase0 + i, phase0 = 1 after a global scan, else 0) gives every particle
// k[i][P][3] and s01[i][P][2]; k_converge runs once per phase.
int thx_pf_converge(thx_ctx* c, const double* k, const double* s01, int nPh, int P,
                    int afterGlobal, int* nPhase) {
    clear_error();
    THX_CHECK(c && k && s01 && nPhase && nPh >= 1 && P >= 1, THX_ERR_ARG,
              "pf_converge: bad argument");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int phase0 = afterGlobal ? 1 : 0, minPhase = afterGlobal ? 10 : 3;
    std::vector<double> s3(3 * (size_t)P * nPh, 0.0), best(2 * (size_t)P, 1.7976931348623157e308);
    for (size_t i = 0; i < (size_t)P * nPh; i++) {
        s3[3 * i] = s01[2 * i];
        s3[3 * i + 1] = s01[2 * i + 1];
    }
    const std::vector<int> ones(P, 1), none(P, -1);
    DevBuf dk, ds, da, db, dn;
    int s = THX_OK;
    if ((s = upload(dk, k, 3 * (size_t)P * nPh, st)) || (s = upload(ds, s3.data(), s3.size(), st)) ||
        (s = upload(da, ones.data(), P, st)) || (s = upload(db, best.data(), best.size(), st)) ||
        (s = upload(dn, none.data(), P, st)))
        return s;
    for (int i = 0; i < nPh; i++)
        hipLaunchKernelGGL(k_converge, dim3(nb(P)), dim3(256), 0, st,
                           dk.as<const double>() + 3 * (size_t)P * i,
                           ds.as<const double>() + 3 * (size_t)P * i, da.as<int>(),
                           db.as<double>(), dn.as<int>(), P, phase_host((uint32_t)(phase0 + i)),
                           minPhase);
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    THX_HIP(hipMemcpy(nPhase, dn.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    DevBuf* all[] = {&dk, &ds, &da, &db, &dn};
    for (DevBuf* b : all) b->release();
    return THX_OK;
}

// The 4x4 helpers of the inferACG stages on n given matrices A[n][16]
// (checking): inv[n][16] and det[n] (zero inverse for a singular matrix),
// top[n][4] the unit eigenvector of the largest eigenvalue of a symmetric A
// with its largest-magnitude entry positive.
int thx_pf_mat4(thx_ctx* c, const double* A, int n, double* inv, double* det, double* top) {
    clear_error();
    THX_CHECK(c && A && inv && det && top && n >= 1, THX_ERR_ARG, "pf_mat4: bad argument");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dA, di, dd, dt;
    int s = THX_OK;
    if ((s = upload(dA, A, 16 * (size_t)n, st)) || (s = di.ensure(sizeof(double) * 16 * n)) ||
        (s = dd.ensure(sizeof(double) * n)) || (s = dt.ensure(sizeof(double) * 4 * n)))
        return s;
    hipLaunchKernelGGL(k_mat4_check, dim3(nb(n)), dim3(256), 0, st, dA.as<const double>(), n,
                       di.as<double>(), dd.as<double>(), dt.as<double>());
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    THX_HIP(hipMemcpy(inv, di.p, sizeof(double) * 16 * n, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(det, dd.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(top, dt.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
    DevBuf* all[] = {&dA, &di, &dd, &dt};
    for (DevBuf* b : all) b->release();
    return THX_OK;
}

// Peak keeping of P sets of n values u (in place): factor[P] in/out
// (setFactor: computed), top[P] the first maximum.
int thx_pf_keep_peak(thx_ctx* c, double* u, int n, int P, double* factor, int setFactor,
                     int* top) {
    clear_error();
    THX_CHECK(c && u && factor && n >= 1 && P >= 1, THX_ERR_ARG, "pf_keep_peak: bad argument");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf du, df, dtop;
    const size_t N = (size_t)P * n;
    int s = THX_OK;
    if ((s = upload(du, u, N, st)) || (s = upload(df, factor, P, st)) ||
        (s = dtop.ensure(sizeof(int) * P)))
        return s;
    hipLaunchKernelGGL(k_keep_peak, dim3(P), dim3(256), 0, st, du.as<double>(), n, nullptr,
                       df.as<double>(), setFactor, 1, nullptr, 0, nullptr, dtop.as<int>());
    THX_HIP(hipGetLastError());
    THX_HIP(hipStreamSynchronize(st));
    THX_HIP(hipMemcpy(u, du.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(factor, df.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    if (top) THX_HIP(hipMemcpy(top, dtop.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    DevBuf* all[] = {&du, &df, &dtop};
    for (DevBuf* b : all) b->release();
    return THX_OK;
}

}  // extern "C"
