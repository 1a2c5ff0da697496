// Particle-filter kernels of the expectation step: weight normalisation,
// systematic resampling with FP64 cumulative sums, effective sample size,
// angular-central-Gaussian (ACG) rotation sampling and perturbation, and
// Gaussian translation sampling.  All batched over particles with ragged
// per-particle sample counts (offsets / counts arrays).
//
// Random numbers: Philox4x32-10 counter streams keyed by (seed, particle,
// sample, stream), so results do not depend on launch shape.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include <hipcub/device/device_radix_sort.hpp>

#include "thx_internal.hpp"

namespace thx {

// ---------------------------------------------------------------------------
// Rotation perturbation (Particle::perturb(pf, PAR_R), src/Particle.cpp:
// 1133-1273, 3-D branch): d ~ ACG(1, pf^2 k1, pf^2 k2, pf^2 k3) and
// q <- mean * d * conj(mean) * q.  One thread per sample.
// ---------------------------------------------------------------------------
__global__ void k_perturb_rot(double* __restrict__ quat, const int* __restrict__ off,
                              const int* __restrict__ cnt, const double* __restrict__ mean,
                              const double* __restrict__ kk, double kmax, PhaseScale pfs, int P,
                              uint64_t seed, PhaseArg pha, const double* __restrict__ dGiven) {
    const int p = blockIdx.x;   // particles on x, the grid dimension with the largest limit
    const int j = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= P || j >= cnt[p]) return;
    const double pf = pfs.get();
    const uint32_t phase = pha.get();
    const size_t i = (size_t)off[p] + j;
    double d[4];
    if (dGiven) {
        for (int a = 0; a < 4; a++) d[a] = dGiven[4 * i + a];
    } else {
        double g[4];
        normals4(seed, (uint32_t)i, phase, kStreamPose + 16, g);
        // k clamped to kmax when kmax > 0 (PERTURB_K_MAX, include/Particle.h:56)
        const double s = pf * pf;
        double k3[3] = {kk[3 * p], kk[3 * p + 1], kk[3 * p + 2]};
        if (kmax > 0)
            for (int a = 0; a < 3; a++) k3[a] = fmin(kmax, k3[a]);
        acg_identity(g, s * k3[0], s * k3[1], s * k3[2], d);
    }
    const double m[4] = {mean[4 * p], mean[4 * p + 1], mean[4 * p + 2], mean[4 * p + 3]};
    const double mc[4] = {m[0], -m[1], -m[2], -m[3]};
    double q[4] = {quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
    double t1[4], t2[4], t3[4];
    qmul(mc, q, t1);      // conj(mean) q
    qmul(d, t1, t2);      // d conj(mean) q
    qmul(m, t2, t3);      // mean d conj(mean) q
    for (int a = 0; a < 4; a++) quat[4 * i + a] = t3[a];
}

// Translation perturbation (Particle::perturb(pf, PAR_T)): bivariate Gaussian
// (gsl_ran_bivariate_gaussian(s0, s1, rho)) scaled by pf, added to t.
__global__ void k_perturb_trans(double* __restrict__ trans, const int* __restrict__ off,
                                const int* __restrict__ cnt, const double* __restrict__ s01rho,
                                double pf, int P, uint64_t seed, uint32_t phase) {
    const int p = blockIdx.x;   // particles on x, the grid dimension with the largest limit
    const int j = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= P || j >= cnt[p]) return;
    const size_t i = (size_t)off[p] + j;
    double g[4];
    normals4(seed, (uint32_t)i, phase, kStreamShift + 16, g);
    const double s0 = s01rho[3 * p], s1 = s01rho[3 * p + 1], rho = s01rho[3 * p + 2];
    const double x = s0 * g[0];
    const double y = s1 * (rho * g[0] + sqrt(fmax(0.0, 1.0 - rho * rho)) * g[1]);
    trans[2 * i] += x * pf;
    trans[2 * i + 1] += y * pf;
}

// ---------------------------------------------------------------------------
// Mixture proposal sampling: particle p has C ACG components, each a mean
// quaternion m_c, concentration k_c in (0, 1] (ACG(diag(1, k, k, k)) rotated
// to m_c) and weight pi_c; plus a Gaussian over shifts (mu, sigma).  Writes
// nRot rotations and nTr translations per particle and the proposal density
// of each rotation, q(x) = sum_c pi_c ACG_c(x) with
// ACG_c(x) = k_c^{-3/2} ((m_c.x)^2 + (1 - (m_c.x)^2) / k_c)^{-2} / (2 pi^2)
// (the normalised form of pdfACG, src/Geometry/DirectionalStat.cpp:13-31,
// on S^3 with q and -q identified).
// ---------------------------------------------------------------------------
__device__ inline double acg_pdf(const double* m, double k, const double* x) {
    const double c = m[0] * x[0] + m[1] * x[1] + m[2] * x[2] + m[3] * x[3];
    const double c2 = c * c;
    const double s = c2 + (1.0 - c2) / k;
    return pow(k, -1.5) / (s * s) / (2.0 * M_PI * M_PI);
}

__global__ void k_sample_mixture(const double* __restrict__ mix, int C, int P, int nRot,
                                 int nTr, uint64_t seed, uint32_t phase,
                                 double* __restrict__ quat, double* __restrict__ qdens,
                                 double* __restrict__ trans) {
    // mix per particle: C x (m0 m1 m2 m3 k pi) then (mu_x mu_y sigma)
    const int p = blockIdx.x;   // particles on x, the grid dimension with the largest limit
    const int j = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double* mp = mix + (size_t)p * (6 * C + 3);
    if (j < nRot) {
        uint32_t w[4];
        Philox::draw(seed, (uint32_t)(p * nRot + j), phase, kStreamPose + 32, w);
        const double u = Philox::u01(w[0], w[1]);
        double acc = 0.0;
        int c = C - 1;
        for (int i = 0; i < C; i++) {
            acc += mp[6 * i + 5];
            if (u <= acc) { c = i; break; }
        }
        double g[4], d[4], q[4];
        normals4(seed, (uint32_t)(p * nRot + j), phase, kStreamPose + 48, g);
        const double k = mp[6 * c + 4];
        acg_identity(g, k, k, k, d);
        qmul(mp + 6 * c, d, q);
        const size_t o = (size_t)p * nRot + j;
        for (int a = 0; a < 4; a++) quat[4 * o + a] = q[a];
        double dens = 0.0;
        for (int i = 0; i < C; i++) dens += mp[6 * i + 5] * acg_pdf(mp + 6 * i, mp[6 * i + 4], q);
        qdens[o] = dens;
    }
    if (j < nTr) {
        double g[4];
        normals4(seed, (uint32_t)(p * nTr + j), phase, kStreamShift + 32, g);
        const double* sh = mp + 6 * C;
        const size_t o = (size_t)p * nTr + j;
        trans[2 * o] = sh[0] + sh[2] * g[0];
        trans[2 * o + 1] = sh[1] + sh[2] * g[1];
    }
}

// ---------------------------------------------------------------------------
// Systematic resampling (Particle::resample(n, pt), src/Particle.cpp:
// 1275-1462), one block per particle: W = w (.) u normalised, FP64 inclusive
// cumulative sum, cdf /= cdf[last], u_j = u0 + j / nOut, idx_j = first i with
// cdf[i] >= u_j, new weight 1 / u_idx (PARTICLE_PRIOR_ONE) normalised.
// u0 is uniform in [0, 1/nOut) from Philox, or given (u0Given >= 0).
// ---------------------------------------------------------------------------
constexpr int kResampleMax = 8192;

__global__ __launch_bounds__(256) void k_resample(
    const double* __restrict__ w, const double* __restrict__ u, const int* __restrict__ offIn,
    const int* __restrict__ nIn, const int* __restrict__ offOut, const int* __restrict__ nOut,
    int P, uint64_t seed, PhaseArg pha, const double* __restrict__ u0Given,
    int* __restrict__ idx, double* __restrict__ wNew, double* __restrict__ essOut, int wShared) {
    const int p = blockIdx.x;
    if (p >= P) return;
    const uint32_t phase = pha.get();
    const int n = nIn[p], m = nOut[p];
    if (n <= 0 || m <= 0 || n > kResampleMax) {   // no samples, or more than the LDS holds
        if (threadIdx.x == 0 && essOut) essOut[p] = -1.0;
        return;
    }
    const double* wp = wShared ? w : w + offIn[p];   // wShared: one prior for every particle
    const double* up = u + offIn[p];
    __shared__ double cdf[kResampleMax];
    __shared__ double part[256];
    // per-thread contiguous ranges, then a scan of the 256 range sums
    const int per = (n + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per, b1 = min(n, b0 + per);
    double s = 0.0, s2 = 0.0;
    for (int i = b0; i < b1; i++) {
        const double v = wp[i] * up[i];
        s += v;
        cdf[i] = s;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) {
            const double v = part[t];
            part[t] = acc;
            acc += v;
        }
    }
    __syncthreads();
    const double offs = part[threadIdx.x];
    for (int i = b0; i < b1; i++) cdf[i] += offs;
    __syncthreads();
    const double total = cdf[n - 1];
    // effective sample size of the normalised weights W
    for (int i = b0; i < b1; i++) {
        const double v = wp[i] * up[i] / total;
        s2 += v * v;
    }
    part[threadIdx.x] = s2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && essOut) essOut[p] = 1.0 / part[0];
    double u0;
    if (u0Given) {
        u0 = u0Given[p];
    } else {
        uint32_t r[4];
        Philox::draw(seed, (uint32_t)p, phase, kStreamPose + 64, r);
        u0 = Philox::u01(r[0], r[1]) / m;
    }
    __syncthreads();
    double ws = 0.0;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        const double uj = (u0 + (double)j / m) * total;   // cdf not normalised: scale u
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cdf[mid] >= uj) hi = mid; else lo = mid + 1;
        }
        idx[offOut[p] + j] = lo;
        const double wn = 1.0 / up[lo];
        wNew[offOut[p] + j] = wn;
        ws += wn;
    }
    part[threadIdx.x] = ws;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    const double tot = part[0];
    for (int j = threadIdx.x; j < m; j += blockDim.x) wNew[offOut[p] + j] /= tot;
}

// ---------------------------------------------------------------------------
// Weighted mean rotation of each particle's samples: principal eigenvector of
// the weighted scatter sum_i w_i q_i q_i^T (the ACG mean that inferACG,
// src/Geometry/DirectionalStat.cpp:87-244, estimates), by power iteration.
// Also returns the eigenvalue ratios lambda_a / lambda_0 (a = 1..3) used as
// the concentration k1..k3 of the perturbation (Particle::calVari,
// src/Particle.cpp:1002-1126).
// ---------------------------------------------------------------------------
__global__ void k_rot_mean(const double* __restrict__ quat, const double* __restrict__ w,
                           const int* __restrict__ off, const int* __restrict__ cnt, int P,
                           double* __restrict__ mean, double* __restrict__ kk) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double A[4][4] = {};
    double ws = 0.0;
    for (int j = 0; j < cnt[p]; j++) {
        const size_t i = (size_t)off[p] + j;
        const double wi = w ? w[i] : 1.0;
        const double* q = quat + 4 * i;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) A[a][b] += wi * q[a] * q[b];
        ws += wi;
    }
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) A[a][b] /= ws;
    // eigen-decomposition of the symmetric 4x4 by cyclic Jacobi
    double Vv[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    for (int sweep = 0; sweep < 30; sweep++) {
        double off2 = 0.0;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) off2 += A[a][b] * A[a][b];
        if (off2 < 1e-30) break;
        for (int a = 0; a < 4; a++)
            for (int b = a + 1; b < 4; b++) {
                if (fabs(A[a][b]) < 1e-300) continue;
                const double th = (A[b][b] - A[a][a]) / (2 * A[a][b]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; k++) {
                    const double akp = A[k][a], akq = A[k][b];
                    A[k][a] = c * akp - s * akq;
                    A[k][b] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; k++) {
                    const double apk = A[a][k], aqk = A[b][k];
                    A[a][k] = c * apk - s * aqk;
                    A[b][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; k++) {
                    const double vkp = Vv[k][a], vkq = Vv[k][b];
                    Vv[k][a] = c * vkp - s * vkq;
                    Vv[k][b] = s * vkp + c * vkq;
                }
            }
    }
    int order[4] = {0, 1, 2, 3};
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++)
            if (A[order[b]][order[b]] > A[order[a]][order[a]]) {
                const int t = order[a]; order[a] = order[b]; order[b] = t;
            }
    const int top = order[0];
    double sgn = Vv[0][top] < 0 ? -1.0 : 1.0;   // canonical sign: w >= 0
    for (int a = 0; a < 4; a++) mean[4 * p + a] = sgn * Vv[a][top];
    const double l0 = fmax(A[top][top], 1e-300);
    for (int a = 1; a < 4; a++) kk[3 * p + a - 1] = fmax(A[order[a]][order[a]], 0.0) / l0;
}

// ---------------------------------------------------------------------------
// Grid-restricted proposal: per particle, the log mixture density of every
// grid rotation and the log Gaussian density of every grid translation; the
// nSel highest of each are selected (threshold by bisection on the value,
// ties by index order).  One block per particle.
// ---------------------------------------------------------------------------
__device__ inline float block_count_ge(const float* v, int n, float thr, int* red) {
    int c = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) c += v[i] >= thr;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += red[w];
    __syncthreads();
    return (float)r;
}

// block_count_ge over values held in registers, with one barrier: the wave
// counts go to slot `par` of red (red[16 par + wave]), and consecutive calls
// alternate the slot, so a slot is rewritten only after every thread has
// passed the barrier of the call in between (and so finished reading it)
template <int NPT>
__device__ inline int block_count_ge_reg(const float (&vr)[NPT], float thr, int* red, int par) {
    int c = 0;
#pragma unroll
    for (int j = 0; j < NPT; j++) c += vr[j] >= thr;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) red[16 * par + (threadIdx.x >> 6)] = c;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += red[16 * par + w];
    return r;
}

// exclusive prefix sum of x over the block in thread order, and its total
// (red: one int per wave, not in use by any thread in flight)
__device__ inline int block_excl_scan(int x, int* red, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        const int t = red[w];
        if (w < wave) before += t;
        total += t;
    }
    return before + inc - x;
}

// select the nSel largest of v[0..n) into out (indices) and outv (values):
// threshold thr by bisection (count(v >= thr) >= nSel, thr as large as
// possible); the output lists the values above the bisection's upper bound
// first, then those in [thr, upper bound] until nSel are taken, each part in
// index order.
// NPT > 0 (n <= NPT blockDim): thread t holds the contiguous run
// [t m, t m + m) of v in registers (m = ceil(n / blockDim)), so thread order is
// index order and the ordered compaction is one block-wide prefix sum.
// NPT == 0: counts and compaction read LDS, 256 values per step.
template <int NPT>
__device__ void block_topk(const float* v, int n, int nSel, int* out, float* outv, float lo,
                           float hi, int* red, int* cnt) {
    if constexpr (NPT > 0) {
        const int m = (n + (int)blockDim.x - 1) / (int)blockDim.x;
        const int i0 = (int)threadIdx.x * m;
        float vr[NPT];
#pragma unroll
        for (int j = 0; j < NPT; j++) vr[j] = (j < m && i0 + j < n) ? v[i0 + j] : -INFINITY;
        // the count is block-uniform, so the early exit is too
        for (int it = 0; it < 40; it++) {
            const float mid = 0.5f * (lo + hi);
            const int cnt0 = block_count_ge_reg<NPT>(vr, mid, red, it & 1);
            if (cnt0 >= nSel) lo = mid; else hi = mid;
            if (cnt0 == nSel) break;
        }
        const float thr = lo;
        int c0 = 0, c1 = 0;
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            c0 += vr[j] > hi;
            c1 += vr[j] >= thr && vr[j] <= hi;
        }
        // both counts in one scan: each stays below 2^16 (n <= 8192)
        int total;
        const int before = block_excl_scan(c0 | (c1 << 16), red + 32, total);
        int o0 = before & 0xffff, o1 = (total & 0xffff) + (before >> 16);
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const float x = vr[j];
            if (x > hi) {
                if (o0 < nSel) { out[o0] = i0 + j; outv[o0] = x; }
                o0++;
            } else if (x >= thr && x <= hi) {
                if (o1 < nSel) { out[o1] = i0 + j; outv[o1] = x; }
                o1++;
            }
        }
        __syncthreads();   // red is free again for the caller
    } else {
        for (int it = 0; it < 40; it++) {
            const float mid = 0.5f * (lo + hi);
            const float cnt0 = block_count_ge(v, n, mid, red);
            if (cnt0 >= nSel) lo = mid; else hi = mid;
            if (cnt0 == nSel) break;
        }
        const float thr = lo;
        if (threadIdx.x == 0) *cnt = 0;
        __syncthreads();
        for (int pass = 0; pass < 2; pass++) {
            for (int base = 0; base < n; base += blockDim.x) {
                const int i = base + threadIdx.x;
                bool take = false;
                if (i < n) take = pass == 0 ? v[i] > hi : (v[i] >= thr && v[i] <= hi);
                // ordered compaction within the block
                const unsigned long long msk = __ballot(take);
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                const int before = __popcll(msk & ((1ull << lane) - 1));
                red[wave] = __popcll(msk);
                __syncthreads();
                int off = *cnt;
                for (int w = 0; w < wave; w++) off += red[w];
                if (take && off + before < nSel) {
                    out[off + before] = i;
                    outv[off + before] = v[i];
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    int tot = 0;
                    for (int w = 0; w < (int)(blockDim.x >> 6); w++) tot += red[w];
                    *cnt = min(nSel, *cnt + tot);
                }
                __syncthreads();
            }
        }
    }
}

constexpr int kSelMaxR = 16384;

// grid quaternions in FP32, one 16-byte load each in the selection kernel
__global__ void k_quat_to_f4(const double* __restrict__ q, int R, float4* __restrict__ qf) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R)
        qf[r] = make_float4((float)q[4 * r], (float)q[4 * r + 1], (float)q[4 * r + 2],
                            (float)q[4 * r + 3]);
}

template <int NPT>
__global__ __launch_bounds__(256) void k_select_grid(
    const double* __restrict__ mix, int C, const float4* __restrict__ gridQ, int R,
    const double* __restrict__ gridT, int T, int nSelR, int nSelT, int* __restrict__ ridx,
    float* __restrict__ rlogq, int* __restrict__ tidx, float* __restrict__ spread) {
    const int p = blockIdx.x;
    const double* mp = mix + (size_t)p * (6 * C + 3);
    extern __shared__ float v[];   // [R]: sized at launch
    __shared__ int red[256];
    __shared__ int cnt;
    __shared__ float tv[512];
    // per component (FP32): mean m, and the density pi k^{-3/2} / (2 pi^2) /
    // (a c^2 + b)^2 with c = m.q, a = 1 - 1/k, b = 1/k
    __shared__ float cm[16][4], ccoef[16], ca[16], cb[16];
    if ((int)threadIdx.x < C) {
        const int c = threadIdx.x;
        const double k = mp[6 * c + 4];
        for (int a = 0; a < 4; a++) cm[c][a] = (float)mp[6 * c + a];
        ccoef[c] = (float)(mp[6 * c + 5] * pow(k, -1.5) / (2.0 * M_PI * M_PI));
        ca[c] = (float)(1.0 - 1.0 / k);
        cb[c] = (float)(1.0 / k);
    }
    __syncthreads();
    float mx = -INFINITY, mn = INFINITY;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const float4 q = gridQ[r];
        float dens = 0.f;
        for (int c = 0; c < C; c++) {
            const float d = cm[c][0] * q.x + cm[c][1] * q.y + cm[c][2] * q.z + cm[c][3] * q.w;
            const float sv = ca[c] * d * d + cb[c];
            // hardware reciprocal (1 ulp): the densities only rank and compare
            dens += ccoef[c] * __builtin_amdgcn_rcpf(sv * sv);
        }
        const float lv = __logf(fmaxf(dens, 1e-37f));
        v[r] = lv;
        mx = fmaxf(mx, lv);
        mn = fminf(mn, lv);
    }
    // block min / max for the bisection range
    __shared__ float rmx[256], rmn[256];
    rmx[threadIdx.x] = mx;
    rmn[threadIdx.x] = mn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            rmx[threadIdx.x] = fmaxf(rmx[threadIdx.x], rmx[threadIdx.x + s]);
            rmn[threadIdx.x] = fminf(rmn[threadIdx.x], rmn[threadIdx.x + s]);
        }
        __syncthreads();
    }
    const float hiR = rmx[0], loR = rmn[0];
    __syncthreads();
    if (spread) {
        // effective number of grid rotations under q: (sum q)^2 / sum q^2
        float s1 = 0.f, s2 = 0.f;
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            const float e = __expf(v[r] - hiR);
            s1 += e;
            s2 += e * e;
        }
        rmx[threadIdx.x] = s1;
        rmn[threadIdx.x] = s2;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                rmx[threadIdx.x] += rmx[threadIdx.x + s];
                rmn[threadIdx.x] += rmn[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) spread[p] = rmx[0] * rmx[0] / rmn[0];
        __syncthreads();
    }
    block_topk<NPT>(v, R, nSelR, ridx + (size_t)p * nSelR, rlogq + (size_t)p * nSelR, loR, hiR, red,
               &cnt);
    // translations: Gaussian log density (the constant does not matter)
    const double* sh = mp + 6 * C;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        const double dx = gridT[2 * t] - sh[0], dy = gridT[2 * t + 1] - sh[1];
        tv[t] = (float)(-(dx * dx + dy * dy) / (2 * sh[2] * sh[2]));
    }
    __syncthreads();
    // exact top-nSelT by rank (value descending, ties by index), listed in
    // index order: T is small, so each thread ranks its translations against
    // all of them instead of bisecting
    int base = 0;
    for (int t0 = 0; t0 < T; t0 += blockDim.x) {
        const int t = t0 + threadIdx.x;
        bool take = false;
        if (t < T) {
            const float x = tv[t];
            int rank = 0;
            for (int u = 0; u < T; u++) {
                const float y = tv[u];
                rank += (y > x) || (y == x && u < t);
            }
            take = rank < nSelT;
        }
        const unsigned long long m = __ballot(take);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        red[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; w++) off += red[w];
        if (take) tidx[(size_t)p * nSelT + off + __popcll(m & ((1ull << lane) - 1))] = t;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) base += red[w];
        __syncthreads();
    }
}

// Sort key of each particle for the selected-pose scorer: the Morton code of
// the mean of its proposal's dominant component, the one with the highest
// density at its mean (pi_c k_c^{-3/2}: the selection's component density
// pi k^{-3/2} / (2 pi^2) / ((1 - 1/k) (m.q)^2 + 1/k)^2 at m.q = +-1).
// Particles with nearby keys select overlapping grid rows.
__global__ void k_sel_keys(const double* __restrict__ mix, int C, int P,
                           uint32_t* __restrict__ key, int* __restrict__ val) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double* mp = mix + (size_t)p * (6 * C + 3);
    int best = 0;
    double bv = -1.0;
    for (int c = 0; c < C; c++) {
        const double k = fmax(mp[6 * c + 4], 1e-300);
        const double v = mp[6 * c + 5] / (k * sqrt(k));
        if (v > bv) { bv = v; best = c; }
    }
    key[p] = quat_morton_dev(mp + 6 * best);
    val[p] = p;
}

__global__ void k_pack_learned(const double* __restrict__ score, const double* __restrict__ ess,
                               const int* __restrict__ bestR, const int* __restrict__ bestT,
                               const float* __restrict__ spread, const float* __restrict__ tail,
                               int P, char* __restrict__ dst) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double* d = reinterpret_cast<double*>(dst);
    d[p] = score[p];
    d[P + p] = ess[p];
    int* i = reinterpret_cast<int*>(d + 2 * (size_t)P);
    i[p] = bestR[p];
    i[P + p] = bestT[p];
    float* f = reinterpret_cast<float*>(i + 2 * (size_t)P);
    f[p] = spread[p];
    f[P + p] = tail[p];
}

int launch_pack_learned(hipStream_t st, const double* score, const double* ess, const int* bestR,
                        const int* bestT, const float* spread, const float* tail, int P,
                        void* dst) {
    hipLaunchKernelGGL(k_pack_learned, dim3((P + 255) / 256), dim3(256), 0, st, score, ess, bestR,
                       bestT, spread, tail, P, static_cast<char*>(dst));
    THX_HIP(hipGetLastError());
    return THX_OK;
}

__global__ void k_gather_rows(const double* __restrict__ src, int rowLen,
                              const int* __restrict__ idx, int n, double* __restrict__ dst) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * rowLen) return;
    const size_t i = e / rowLen, j = e % rowLen;
    dst[e] = src[(size_t)idx[i] * rowLen + j];
}

int launch_gather_rows(hipStream_t st, const double* src, int rowLen, const int* idx, int n,
                       double* dst) {
    if (n <= 0) return THX_OK;
    const size_t m = (size_t)n * rowLen;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, src,
                       rowLen, idx, n, dst);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_select_order(hipStream_t st, const double* mix, int C, int P, uint32_t* keyA,
                    uint32_t* keyB, int* val, int* order, void* tmp, size_t* tmpBytes) {
    if (!tmp) {
        THX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, *tmpBytes, keyA, keyB, val, order, P,
                                                   0, 30, st));
        return THX_OK;
    }
    hipLaunchKernelGGL(k_sel_keys, dim3((P + 255) / 256), dim3(256), 0, st, mix, C, P, keyA, val);
    THX_HIP(hipGetLastError());
    // stable LSD radix sort over the 30 key bits: equal keys keep particle order
    THX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, *tmpBytes, keyA, keyB, val, order, P, 0, 30,
                                               st));
    return THX_OK;
}

int pf_select_grid(hipStream_t st, const double* mix, int C, int P, const double* gridQ, int R,
                   const double* gridT, int T, int nSelR, int nSelT, int* ridx, float* rlogq,
                   int* tidx, float* spread, float4* gridQf) {
    THX_CHECK(C >= 1 && C <= 16 && R <= kSelMaxR && T <= 512 && nSelR <= R && nSelT <= T,
              THX_ERR_RANGE, "select_grid: sizes out of range");
    hipLaunchKernelGGL(k_quat_to_f4, dim3((R + 255) / 256), dim3(256), 0, st, gridQ, R, gridQf);
    // grids up to 8,192 rotations count in registers (32 per thread at most)
    auto launch = [&](auto kern) -> int {
        if (sizeof(float) * R > 32768)
            THX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(sizeof(float) * R)));
        hipLaunchKernelGGL(kern, dim3(P), dim3(256), sizeof(float) * R, st, mix, C, gridQf, R,
                           gridT, T, nSelR, nSelT, ridx, rlogq, tidx, spread);
        THX_HIP(hipGetLastError());
        return THX_OK;
    };
    if (R <= 256 * 8) return launch(&k_select_grid<8>);
    if (R <= 256 * 16) return launch(&k_select_grid<16>);
    if (R <= 256 * 24) return launch(&k_select_grid<24>);
    if (R <= 256 * 32) return launch(&k_select_grid<32>);
    return launch(&k_select_grid<0>);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int pf_perturb_rot(hipStream_t st, double* quat, const int* off, const int* cnt, int maxCnt,
                   const double* mean, const double* kk, double kmax, double pf, int P,
                   uint64_t seed, uint32_t phase, const double* dGiven) {
    dim3 grid(P, (maxCnt + 255) / 256);
    hipLaunchKernelGGL(k_perturb_rot, grid, dim3(256), 0, st, quat, off, cnt, mean, kk, kmax,
                       scale_host(pf), P, seed, phase_host(phase), dGiven);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// the same with the perturbation scale and the random-stream phase read from
// device memory by the kernel (a graph replayed over the local-search phases)
int pf_perturb_rot(hipStream_t st, double* quat, const int* off, const int* cnt, int maxCnt,
                   const double* mean, const double* kk, double kmax, PhaseScale pf, int P,
                   uint64_t seed, PhaseArg phase) {
    dim3 grid(P, (maxCnt + 255) / 256);
    hipLaunchKernelGGL(k_perturb_rot, grid, dim3(256), 0, st, quat, off, cnt, mean, kk, kmax, pf,
                       P, seed, phase, (const double*)nullptr);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_perturb_trans(hipStream_t st, double* trans, const int* off, const int* cnt, int maxCnt,
                     const double* s01rho, double pf, int P, uint64_t seed, uint32_t phase) {
    dim3 grid(P, (maxCnt + 255) / 256);
    hipLaunchKernelGGL(k_perturb_trans, grid, dim3(256), 0, st, trans, off, cnt, s01rho, pf, P,
                       seed, phase);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_sample_mixture(hipStream_t st, const double* mix, int C, int P, int nRot, int nTr,
                      uint64_t seed, uint32_t phase, double* quat, double* qdens,
                      double* trans) {
    const int n = nRot > nTr ? nRot : nTr;
    dim3 grid(P, (n + 255) / 256);
    hipLaunchKernelGGL(k_sample_mixture, grid, dim3(256), 0, st, mix, C, P, nRot, nTr, seed, phase,
                       quat, qdens, trans);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_resample(hipStream_t st, const double* w, const double* u, const int* offIn,
                const int* nIn, const int* offOut, const int* nOut, int P, uint64_t seed,
                uint32_t phase, const double* u0Given, int* idx, double* wNew, double* ess,
                int wShared) {
    hipLaunchKernelGGL(k_resample, dim3(P), dim3(256), 0, st, w, u, offIn, nIn, offOut, nOut, P,
                       seed, phase_host(phase), u0Given, idx, wNew, ess, wShared);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_resample(hipStream_t st, const double* w, const double* u, const int* offIn,
                const int* nIn, const int* offOut, const int* nOut, int P, uint64_t seed,
                PhaseArg phase, int* idx, double* wNew) {
    hipLaunchKernelGGL(k_resample, dim3(P), dim3(256), 0, st, w, u, offIn, nIn, offOut, nOut, P,
                       seed, phase, (const double*)nullptr, idx, wNew, (double*)nullptr, 0);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int pf_rot_mean(hipStream_t st, const double* quat, const double* w, const int* off,
                const int* cnt, int P, double* mean, double* kk) {
    hipLaunchKernelGGL(k_rot_mean, dim3((P + 127) / 128), dim3(128), 0, st, quat, w, off, cnt, P,
                       mean, kk);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

}  // namespace thx
