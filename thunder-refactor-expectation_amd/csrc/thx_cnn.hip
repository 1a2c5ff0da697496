// Learned pose proposal: a small CNN that maps a particle image to an
// angular-central-Gaussian (ACG) mixture over rotations and a Gaussian over
// shifts, with hand-written forward, backward and Adam kernels (FP32).
//
// Input: the particle's Fourier coefficients with 0 <= iCol < 32 and
// -32 <= iRow < 32 (the half plane of the pixel set; each stands for itself
// and its conjugate), phase-flipped by sign(CTF) and inverse-transformed to a
// 64 x 64 real image, normalised to zero mean and unit variance.
// Network (NCHW):
//   conv 1->16 5x5/2 (64->32) ReLU, conv 16->32 3x3/2 (->16) ReLU,
//   conv 32->64 3x3/2 (->8) ReLU, conv 64->128 3x3/2 (->4) ReLU,
//   fc 2048->384 ReLU, fc 384->NOUT (NOUT = 6 C + 3).
// Output per component c: raw quaternion (4) -> m_c = raw / |raw|,
//   a -> k_c = exp(-(1 + 6 sigmoid(a))), logit -> pi_c = softmax;
// then shift mean (2) and b -> sigma = exp(b) clipped to [0.2, 20].
// Loss: -log sum_c pi_c ACG_c(q) - log N(t; mu, sigma^2 I), with
//   ACG_c(q) = k^{-3/2} ((m.q)^2 + (1 - (m.q)^2) / k)^{-2} / (2 pi^2).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "thx_internal.hpp"

namespace thx {

constexpr int kImg = 64;
constexpr int kMaxBatch = 512;     // training minibatch limit
constexpr int kInferBatch = 16384;  // images per inference pass

struct ConvShape {
    int ci, hi, co, ho, ks, st, pd;
};
static const ConvShape kConv[4] = {
    {1, 64, 16, 32, 5, 2, 2}, {16, 32, 32, 16, 3, 2, 1}, {32, 16, 64, 8, 3, 2, 1},
    {64, 8, 128, 4, 3, 2, 1}};
constexpr int kFcIn = 128 * 4 * 4, kFcHid = 384;

// ---------------------------------------------------------------------------
// Images: 64 x 64 inverse DFT of the phase-flipped coefficients with
// 0 <= iCol < 32 and -32 <= iRow < 32 (each standing for itself and its
// Hermitian partner), then normalisation.  One wave per particle, both
// passes of the separable transform as real products on FP32 MFMA
// (v_mfma_f32_32x32x2f32) with w(t) = exp(2 pi i t / 64):
//   pass 1  H[j][x] = sum_i G[j][i] w(i x)                (j = row + 32)
//           as [Hr | Hi] (64 x 128) = [Gr | Gi] (64 x 64) B1, with
//           B1 = [[cos, sin], [-sin, cos]] of (i x), G staged in LDS;
//   pass 2  img[y][x] = Re sum_j H[j][x] w((j - 32) y)
//           as img (64 x 64) = [cos | -sin] of ((j - 32) y) (64 x 128) [Hr; Hi],
// where pass 2 takes H straight from pass 1's accumulators: its k order is
// the accumulator row order, so no data moves between the passes.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kImgLD = 65;   // LDS row stride of G (floats): conflict-free column reads

__global__ __launch_bounds__(256) void k_images(const float2* __restrict__ dat,
                                                const float* __restrict__ ctf,
                                                const int* __restrict__ iCol,
                                                const int* __restrict__ iRow, int K, int P,
                                                float* __restrict__ img) {
    __shared__ __attribute__((aligned(16))) float G[4][64 * kImgLD];   // per wave: [row j][Re i | Im i]
    __shared__ float2 tw[64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hl = lane >> 5;
    const int p = blockIdx.x * 4 + wave;
    if (tid < 64) {
        float sn, cs;
        sincospif(tid / 32.f, &sn, &cs);
        tw[tid] = make_float2(cs, sn);
    }
    float* g = G[wave];
    for (int e = lane; e < 64 * kImgLD; e += 64) g[e] = 0.f;
    __syncthreads();
    if (p < P) {
        // eight coefficients per lane in flight at a time
        constexpr int U = 8;
        for (int k0 = 0; k0 < K; k0 += 64 * U) {
            float2 x[U];
            float cv[U];
            int ij[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = k0 + 64 * u + lane;
                ij[u] = -1;
                if (k < K) {
                    const int i = iCol[k], j = iRow[k];
                    if (i < 32 && j >= -32 && j < 32 && !(i == 0 && j == 0)) {
                        ij[u] = (j + 32) * kImgLD + i;
                        x[u] = dat[(size_t)p * K + k];
                        cv[u] = ctf[(size_t)p * K + k];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (ij[u] >= 0) {
                    const float f = cv[u] >= 0.f ? 2.f : -2.f;
                    g[ij[u]] = f * x[u].x;
                    g[ij[u] + 32] = f * x[u].y;
                }
        }
    }
    __syncthreads();
    if (p >= P) return;   // no block-wide synchronisation follows

    // ---- pass 1: acc[mt][nt], rows j = 32 mt + .., columns n = 32 nt + l32
    // (nt 0, 1: Hr at x = n; nt 2, 3: Hi at x = n - 64)
    floatx16 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = floatx16{0};
    for (int s = 0; s < 32; s++) {
        const int k = 2 * s + hl;   // this lane's k: Re (k < 32) or Im (k >= 32) of i = k & 31
        const int i = k & 31;
        const float a0 = g[l32 * kImgLD + k];
        const float a1 = g[(32 + l32) * kImgLD + k];
        const float2 w0 = tw[(i * l32) & 63];          // x = l32
        const float2 w1 = tw[(i * (32 + l32)) & 63];   // x = 32 + l32
        // B1[k][n]: Hr columns cos (k < 32) / -sin (k >= 32); Hi columns sin / cos
        const bool re = k < 32;
        const float b0 = re ? w0.x : -w0.y, b1 = re ? w1.x : -w1.y;
        const float b2 = re ? w0.y : w0.x, b3 = re ? w1.y : w1.x;
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[0][2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b2, acc[0][2], 0, 0, 0);
        acc[0][3] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b3, acc[0][3], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        acc[1][2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b2, acc[1][2], 0, 0, 0);
        acc[1][3] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b3, acc[1][3], 0, 0, 0);
    }

    // ---- pass 2, one half of the rows y = 32 my + .. at a time: out[nx],
    // columns x = 32 nx + l32.  Step (mt, e): this lane's k is row
    // j = 32 mt + 8 (e / 4) + 4 hl + e % 4 of pass 1, whose Hr / Hi at
    // x = 32 nx + l32 it already holds in acc[mt][nx] / acc[mt][2 + nx]
    // element e.  The unnormalised image goes to the wave's LDS rows (G is
    // no longer read), 64 floats per row.
    float s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int my = 0; my < 2; my++) {
        floatx16 out[2];
        out[0] = floatx16{0};
        out[1] = floatx16{0};
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int j = 32 * mt + 8 * (e >> 2) + 4 * hl + (e & 3);
                const float2 v = tw[((j - 32) * (32 * my + l32)) & 63];
#pragma unroll
                for (int nx = 0; nx < 2; nx++) {
                    out[nx] = __builtin_amdgcn_mfma_f32_32x32x2f32(v.x, acc[mt][nx][e], out[nx], 0, 0, 0);
                    out[nx] = __builtin_amdgcn_mfma_f32_32x32x2f32(-v.y, acc[mt][2 + nx][e], out[nx], 0, 0, 0);
                }
            }
#pragma unroll
        for (int nx = 0; nx < 2; nx++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const float v = out[nx][e];
                s1 += v;
                s2 += v * v;
                const int y = 32 * my + 8 * (e >> 2) + 4 * hl + (e & 3);
                g[y * 64 + 32 * nx + l32] = v;
            }
    }

    // ---- normalisation over the 4096 pixels of the wave's image
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    const float mean = s1 / 4096.f;
    const float var = fmaxf(s2 / 4096.f - mean * mean, 1e-20f);
    const float inv = rsqrtf(var);
    float4* im = reinterpret_cast<float4*>(img + (size_t)p * 4096);
    const float4* src = reinterpret_cast<const float4*>(g);
#pragma unroll
    for (int q = 0; q < 16; q++) {
        float4 v = src[64 * q + lane];
        v.x = (v.x - mean) * inv;
        v.y = (v.y - mean) * inv;
        v.z = (v.z - mean) * inv;
        v.w = (v.w - mean) * inv;
        im[64 * q + lane] = v;
    }
}

// ---------------------------------------------------------------------------
// The same images for the BF16 proposal passes, both passes on
// v_mfma_f32_32x32x16_bf16 (FP32 accumulation): G, the pass-1 twiddle matrix
// and H are rounded to BF16.  Pass 1 as above, G in BF16 rows (16-byte
// fragment reads) against a block-shared table B1t[n][k]; pass 2 takes H from
// pass 1's accumulators with its k order chosen to match them: step
// (mt, q0, part) covers rows j = 32 mt + 8 q + 4 h + r (q = q0, q0 + 1,
// r = 0..3) of Hr (part 0) or Hi (part 1), lane half h holding
// acc[mt][part nt + ...][4 q0 .. 4 q0 + 7]; the matching A values
// cos / -sin((j - 32) y) come from a block-shared table C2[y][k'] in that
// order.  The image stays in registers through the normalisation.
// ---------------------------------------------------------------------------
typedef __bf16 img_bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int img_u32x4 __attribute__((ext_vector_type(4)));
constexpr int kImgLDB = 72;    // BF16 row stride of G and B1t (64 + 8)
constexpr int kImgLDC = 136;   // BF16 row stride of C2 (128 + 8)

__device__ inline unsigned short img_bf16(float v) {
    return __builtin_bit_cast(unsigned short, (__bf16)v);
}

__global__ __launch_bounds__(256) void k_images_bf16(const float2* __restrict__ dat,
                                                     const float* __restrict__ ctf,
                                                     const int* __restrict__ iCol,
                                                     const int* __restrict__ iRow, int K, int P,
                                                     float* __restrict__ img) {
    __shared__ __attribute__((aligned(16))) unsigned short G[4][64 * kImgLDB];
    __shared__ __attribute__((aligned(16))) unsigned short B1t[128 * kImgLDB];
    __shared__ __attribute__((aligned(16))) unsigned short C2[64 * kImgLDC];
    __shared__ float2 tw[64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hl = lane >> 5;
    const int p = blockIdx.x * 4 + wave;
    if (tid < 64) {
        float sn, cs;
        sincospif(tid / 32.f, &sn, &cs);
        tw[tid] = make_float2(cs, sn);
    }
    unsigned short* g = G[wave];
    for (int e = lane; e < 64 * kImgLDB; e += 64) g[e] = 0;
    __syncthreads();
    // B1t[n][k]: n < 64: Hr at x = n (cos for k < 32, -sin for k >= 32);
    // n >= 64: Hi at x = n - 64 (sin, cos); i = k & 31
    for (int e = tid; e < 128 * 64; e += 256) {
        const int n = e >> 6, k = e & 63, i = k & 31, x = n & 63;
        const float2 w = tw[(i * x) & 63];
        const bool re = k < 32;
        const float v = n < 64 ? (re ? w.x : -w.y) : (re ? w.y : w.x);
        B1t[n * kImgLDB + k] = img_bf16(v);
    }
    // C2[y][k']: k' = 16 step + 8 h + e, step = 4 mt + 2 (q0 / 2) + part,
    // j = 32 mt + 8 (q0 + e / 4) + 4 h + e % 4; cos((j - 32) y) for Hr, -sin for Hi
    for (int e2 = tid; e2 < 64 * 128; e2 += 256) {
        const int y = e2 >> 7, kk = e2 & 127;
        const int step = kk >> 4, h = (kk >> 3) & 1, e = kk & 7;
        const int mt = step >> 2, q0 = 2 * ((step >> 1) & 1), part = step & 1;
        const int j = 32 * mt + 8 * (q0 + (e >> 2)) + 4 * h + (e & 3);
        const float2 w = tw[((j - 32) * y) & 63];
        C2[y * kImgLDC + kk] = img_bf16(part == 0 ? w.x : -w.y);
    }
    if (p < P) {
        constexpr int U = 8;
        for (int k0 = 0; k0 < K; k0 += 64 * U) {
            float2 x[U];
            float cv[U];
            int ij[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = k0 + 64 * u + lane;
                ij[u] = -1;
                if (k < K) {
                    const int i = iCol[k], j = iRow[k];
                    if (i < 32 && j >= -32 && j < 32 && !(i == 0 && j == 0)) {
                        ij[u] = (j + 32) * kImgLDB + i;
                        x[u] = dat[(size_t)p * K + k];
                        cv[u] = ctf[(size_t)p * K + k];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (ij[u] >= 0) {
                    const float f = cv[u] >= 0.f ? 2.f : -2.f;
                    g[ij[u]] = img_bf16(f * x[u].x);
                    g[ij[u] + 32] = img_bf16(f * x[u].y);
                }
        }
    }
    __syncthreads();
    if (p >= P) return;   // no block-wide synchronisation follows

    // ---- pass 1: acc[mt][nt], rows j = 32 mt + .., columns n = 32 nt + l32
    floatx16 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = floatx16{0};
#pragma unroll
    for (int st = 0; st < 4; st++) {
        img_u32x4 fa[2], fb[4];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
            fa[mt] = *reinterpret_cast<const img_u32x4*>(g + (32 * mt + l32) * kImgLDB + 16 * st + 8 * hl);
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
            fb[nt] = *reinterpret_cast<const img_u32x4*>(B1t + (32 * nt + l32) * kImgLDB + 16 * st + 8 * hl);
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(img_bf16x8, fa[mt]), __builtin_bit_cast(img_bf16x8, fb[nt]),
                    acc[mt][nt], 0, 0, 0);
    }

    // ---- pass 2: out[my][nx] rows y = 32 my + .., columns x = 32 nx + l32
    floatx16 out[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) out[a][b] = floatx16{0};
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int qp = 0; qp < 2; qp++)
#pragma unroll
            for (int part = 0; part < 2; part++) {
                const int step = 4 * mt + 2 * qp + part;
                img_u32x4 fa[2];
#pragma unroll
                for (int my = 0; my < 2; my++)
                    fa[my] = *reinterpret_cast<const img_u32x4*>(C2 + (32 * my + l32) * kImgLDC + 16 * step + 8 * hl);
#pragma unroll
                for (int nx = 0; nx < 2; nx++) {
                    const floatx16& hsrc = acc[mt][2 * part + nx];
                    img_u32x4 fbv;
#pragma unroll
                    for (int w = 0; w < 4; w++) {
                        const int e0 = 8 * qp + 2 * w;   // accumulator elements 4 q0 + 2 w, + 1
                        fbv[w] = (unsigned)img_bf16(hsrc[e0]) | ((unsigned)img_bf16(hsrc[e0 + 1]) << 16);
                    }
#pragma unroll
                    for (int my = 0; my < 2; my++)
                        out[my][nx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(img_bf16x8, fa[my]), __builtin_bit_cast(img_bf16x8, fbv),
                            out[my][nx], 0, 0, 0);
                }
            }

    // ---- normalisation over the 4096 pixels of the wave's image
    float s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int my = 0; my < 2; my++)
#pragma unroll
        for (int nx = 0; nx < 2; nx++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const float v = out[my][nx][e];
                s1 += v;
                s2 += v * v;
            }
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    const float mean = s1 / 4096.f;
    const float var = fmaxf(s2 / 4096.f - mean * mean, 1e-20f);
    const float inv = rsqrtf(var);
    float* im = img + (size_t)p * 4096;
#pragma unroll
    for (int my = 0; my < 2; my++)
#pragma unroll
        for (int nx = 0; nx < 2; nx++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int y = 32 * my + 8 * (e >> 2) + 4 * hl + (e & 3);
                im[y * 64 + 32 * nx + l32] = (out[my][nx][e] - mean) * inv;
            }
}

// ---------------------------------------------------------------------------
// Layers as GEMMs on FP32 MFMA.
// C[m][n] = sum_k A(m, k) B(n, k): a block of 256 threads holds WM x WN wave
// tiles of 32 x 32 (v_mfma_f32_32x32x2f32); k advances 16 at a time through
// LDS while the next step's operands load into registers.  A is loaded with
// lanes along k; B with lanes along n (BKF false, for gathers such as
// im2col) or along k (BKF true, for k-contiguous rows).  blockIdx.z splits
// the k range (partial sums go to the store functor, which then adds
// atomically).  Operand functors return A(m, k) / B(n, k) in range only.
// With BF (inference only) each 16-deep step is one v_mfma_f32_32x32x16_bf16
// on operands rounded to BF16 (FP32 accumulation) instead of eight FP32
// MFMAs: lane (m, h) takes k = 8 h .. 8 h + 7 of the step.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <int WM, int WN, bool BKF, class LA, class LB, class ST, bool BF = false>
__global__ __launch_bounds__(256) void k_gemm(LA la, LB lb, ST st, int M, int N, int Kd,
                                              int kPer) {
    static_assert(WM * WN == 4, "four waves per block");
    constexpr int BM = 32 * WM, BN = 32 * WN, BK = 16;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    __shared__ float As[BK][BM + 1];
    __shared__ float Bs[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kb = blockIdx.z * kPer, ke = min(Kd, kb + kPer);
    float ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + 256 * i, k = k0 + (e & 15), m = m0 + (e >> 4);
            ra[i] = (m < M && k < ke) ? la(m, k) : 0.f;
        }
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int e = tid + 256 * i;
            const int k = k0 + (BKF ? (e & 15) : e / BN), n = n0 + (BKF ? (e >> 4) : e % BN);
            rb[i] = (n < N && k < ke) ? lb(n, k) : 0.f;
        }
    };
    floatx16 acc = {0};
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + 256 * i;
            As[e & 15][e >> 4] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int e = tid + 256 * i;
            if constexpr (BKF) Bs[e & 15][e >> 4] = rb[i];
            else Bs[e / BN][e % BN] = rb[i];
        }
        __syncthreads();
        if (k0 + BK < ke) load(k0 + BK);
        if constexpr (BF) {
            bf16x8_t a, b;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                a[e] = (__bf16)As[8 * (lane >> 5) + e][32 * wm + (lane & 31)];
                b[e] = (__bf16)Bs[8 * (lane >> 5) + e][32 * wn + (lane & 31)];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        } else {
#pragma unroll
            for (int s = 0; s < BK; s += 2) {
                const float a = As[s + (lane >> 5)][32 * wm + (lane & 31)];
                const float b = Bs[s + (lane >> 5)][32 * wn + (lane & 31)];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
        }
    }
    const int n = n0 + 32 * wn + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + 32 * wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m < M && n < N) st(m, n, acc[r]);
    }
}

// row-major operand: X(r, k) = p[r * ld + k]
struct RowOp {
    const float* p;
    int ld;
    __device__ float operator()(int r, int k) const { return p[(size_t)r * ld + k]; }
};
// transposed operand: X(r, k) = p[k * ld + r]
struct ColOp {
    const float* p;
    int ld;
    __device__ float operator()(int r, int k) const { return p[(size_t)k * ld + r]; }
};

template <int L> struct Geo;
template <> struct Geo<0> { static constexpr int CI = 1, HI = 64, CO = 16, HO = 32, KS = 5, S = 2, PD = 2; };
template <> struct Geo<1> { static constexpr int CI = 16, HI = 32, CO = 32, HO = 16, KS = 3, S = 2, PD = 1; };
template <> struct Geo<2> { static constexpr int CI = 32, HI = 16, CO = 64, HO = 8, KS = 3, S = 2, PD = 1; };
template <> struct Geo<3> { static constexpr int CI = 64, HI = 8, CO = 128, HO = 4, KS = 3, S = 2, PD = 1; };

// im2col of a conv input: X(n, k) with n = (b, oy, ox), k = (c, ky, kx)
template <int L> struct Im2col {
    using G = Geo<L>;
    const float* in;
    __device__ float operator()(int n, int k) const {
        constexpr int HW = G::HO * G::HO, KK = G::KS * G::KS;
        const int b = n / HW, pos = n % HW, oy = pos / G::HO, ox = pos % G::HO;
        const int c = k / KK, r = k % KK, ky = r / G::KS, kx = r % G::KS;
        const int iy = oy * G::S - G::PD + ky, ix = ox * G::S - G::PD + kx;
        if (iy < 0 || iy >= G::HI || ix < 0 || ix >= G::HI) return 0.f;
        return in[((size_t)(b * G::CI + c) * G::HI + iy) * G::HI + ix];
    }
};
// im2col with the two indices swapped (weight-gradient GEMM: n = k_conv,
// k = (b, pos)), plus a column of ones at n = CI KS KS for the bias gradient
template <int L> struct Im2colT {
    Im2col<L> x;
    __device__ float operator()(int n, int k) const {
        return n == Geo<L>::CI * Geo<L>::KS * Geo<L>::KS ? 1.f : x(k, n);
    }
};
// output gradient as the A operand of the weight gradient: G(co, (b, pos))
template <int L> struct GradRows {
    using G = Geo<L>;
    const float* g;
    __device__ float operator()(int co, int k) const {
        constexpr int HW = G::HO * G::HO;
        const int b = k / HW, pos = k % HW;
        return g[((size_t)b * G::CO + co) * HW + pos];
    }
};
// transposed-conv gather for the input gradient: X(n, k) with
// n = (b, iy, ix) and k = (co, ky, kx); zero where no output tap lands
template <int L> struct GradCols {
    using G = Geo<L>;
    const float* g;
    __device__ float operator()(int n, int k) const {
        constexpr int HWI = G::HI * G::HI, KK = G::KS * G::KS;
        const int b = n / HWI, pos = n % HWI, iy = pos / G::HI, ix = pos % G::HI;
        const int co = k / KK, r = k % KK, ky = r / G::KS, kx = r % G::KS;
        const int ty = iy + G::PD - ky, tx = ix + G::PD - kx;
        if (ty < 0 || tx < 0 || ty % G::S || tx % G::S) return 0.f;
        const int oy = ty / G::S, ox = tx / G::S;
        if (oy >= G::HO || ox >= G::HO) return 0.f;
        return g[((size_t)(b * G::CO + co) * G::HO + oy) * G::HO + ox];
    }
};
// weights as the A operand of the input gradient: W(c, (co, ky, kx))
template <int L> struct WeightT {
    using G = Geo<L>;
    const float* w;
    __device__ float operator()(int c, int k) const {
        constexpr int KK = G::KS * G::KS;
        const int co = k / KK, r = k % KK;
        return w[((size_t)co * G::CI + c) * KK + r];
    }
};

// conv output (m = co, n = (b, pos)) + bias, ReLU
template <int L> struct ConvStore {
    using G = Geo<L>;
    float* out;
    const float* bias;
    __device__ void operator()(int m, int n, float v) const {
        constexpr int HW = G::HO * G::HO;
        const int b = n / HW, pos = n % HW;
        out[((size_t)b * G::CO + m) * HW + pos] = fmaxf(v + bias[m], 0.f);
    }
};
// input gradient store (m = c, n = (b, pos))
template <int L> struct GradInStore {
    using G = Geo<L>;
    float* din;
    __device__ void operator()(int m, int n, float v) const {
        constexpr int HW = G::HI * G::HI;
        const int b = n / HW, pos = n % HW;
        din[((size_t)b * G::CI + m) * HW + pos] = v;
    }
};
// fully-connected output y[m][n] (+ bias[n], optional ReLU)
struct FcStore {
    float* y;
    const float* bias;
    int O, relu;
    __device__ void operator()(int m, int n, float v) const {
        v += bias[n];
        y[(size_t)m * O + n] = relu ? fmaxf(v, 0.f) : v;
    }
};
// plain store y[m][n] = v (ld O)
struct PlainStore {
    float* y;
    int O;
    __device__ void operator()(int m, int n, float v) const { y[(size_t)m * O + n] = v; }
};
// weight and bias gradient partial sums, one slab per k split (blockIdx.z):
// part[z][m][n] with column KC of the GEMM the bias; reduced in split order
// afterwards, so training is deterministic
struct WgradStore {
    float* part;
    int ld, slab;
    __device__ void operator()(int m, int n, float v) const {
        part[(size_t)blockIdx.z * slab + (size_t)m * ld + n] = v;
    }
};

// dw[m][n] (n < KC) and db[m] (n == KC) = sum over splits z of part[z][m][n]
__global__ void k_split_reduce(const float* __restrict__ part, int nz, int CO, int KC,
                               float* __restrict__ dw, float* __restrict__ db) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ld = KC + 1;
    if (i >= CO * ld) return;
    float s = 0.f;
    for (int z = 0; z < nz; z++) s += part[(size_t)z * CO * ld + i];
    const int m = i / ld, n = i % ld;
    if (n == KC) db[m] = s;
    else dw[(size_t)m * KC + n] = s;
}
template <int WM, int WN, bool BKF, class LA, class LB, class ST, bool BF = false>
static int gemm(hipStream_t st, LA la, LB lb, ST sto, int M, int N, int Kd, int splits = 1) {
    constexpr int BM = 32 * WM, BN = 32 * WN;
    const int kPer = ((Kd + splits - 1) / splits + 15) / 16 * 16;
    const int nz = (Kd + kPer - 1) / kPer;
    hipLaunchKernelGGL((k_gemm<WM, WN, BKF, LA, LB, ST, BF>),
                       dim3((N + BN - 1) / BN, (M + BM - 1) / BM, nz), dim3(256), 0, st, la, lb,
                       sto, M, N, Kd, kPer);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// g *= (act > 0): back through a ReLU whose output is act
__global__ void k_relu_back(float* __restrict__ g, const float* __restrict__ act, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && act[i] <= 0.f) g[i] = 0.f;
}

// db[o] = sum_b g[b][o]
__global__ void k_col_sum(const float* __restrict__ g, float* __restrict__ db, int B, int O) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= O) return;
    float s = 0.f;
    for (int bb = 0; bb < B; bb++) s += g[(size_t)bb * O + o];
    db[o] = s;
}

// ---------------------------------------------------------------------------
// Output head: mixture parameters, loss and its gradient w.r.t. the outputs.
// ---------------------------------------------------------------------------
__device__ inline float sigm(float a) { return 1.f / (1.f + expf(-a)); }

// mix (double): C x (m0..m3, k, pi) + (mu_x, mu_y, sigma)
__global__ void k_head(const float* __restrict__ out, int B, int C, double* __restrict__ mix) {
    const int bb = blockIdx.x * blockDim.x + threadIdx.x;
    if (bb >= B) return;
    const int NOUT = 6 * C + 3;
    const float* o = out + (size_t)bb * NOUT;
    double* m = mix + (size_t)bb * NOUT;
    float mx = -INFINITY;
    for (int c = 0; c < C; c++) mx = fmaxf(mx, o[6 * c + 5]);
    double zs = 0.0;
    for (int c = 0; c < C; c++) zs += exp((double)o[6 * c + 5] - mx);
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        for (int a = 0; a < 4; a++) m[6 * c + a] = r[a] / n;
        m[6 * c + 4] = exp(-(1.0 + 6.0 * sigm(r[4])));
        m[6 * c + 5] = exp((double)r[5] - mx) / zs;
    }
    m[6 * C] = o[6 * C];
    m[6 * C + 1] = o[6 * C + 1];
    m[6 * C + 2] = fmin(20.0, fmax(0.2, exp((double)o[6 * C + 2])));
}

// loss[b] and dL/dout[b] for targets q (4, double) and t (2, double); the
// batch loss is the mean, so the gradient carries 1 / B.
__global__ void k_loss(const float* __restrict__ out, const double* __restrict__ qt,
                       const double* __restrict__ tt, int B, int C, float* __restrict__ loss,
                       float* __restrict__ gout) {
    const int bb = blockIdx.x * blockDim.x + threadIdx.x;
    if (bb >= B) return;
    const int NOUT = 6 * C + 3;
    const float* o = out + (size_t)bb * NOUT;
    float* g = gout + (size_t)bb * NOUT;
    const double q[4] = {qt[4 * bb], qt[4 * bb + 1], qt[4 * bb + 2], qt[4 * bb + 3]};
    double lmax = -INFINITY;
    double lc[16], logpi[16], dlogA_dd[16], dlogA_dk[16], dd[16], kc[16];
    double mxl = -INFINITY;
    for (int c = 0; c < C; c++) mxl = fmax(mxl, (double)o[6 * c + 5]);
    double zs = 0.0;
    for (int c = 0; c < C; c++) zs += exp((double)o[6 * c + 5] - mxl);
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        double d = 0.0;
        for (int a = 0; a < 4; a++) d += r[a] / n * q[a];
        const double k = exp(-(1.0 + 6.0 * (double)sigm(r[4])));
        const double s = d * d + (1.0 - d * d) / k;
        const double logA = -log(2.0 * M_PI * M_PI) - 1.5 * log(k) - 2.0 * log(s);
        logpi[c] = (double)r[5] - mxl - log(zs);
        lc[c] = logpi[c] + logA;
        lmax = fmax(lmax, lc[c]);
        dd[c] = d;
        kc[c] = k;
        dlogA_dd[c] = -4.0 * d * (1.0 - 1.0 / k) / s;
        dlogA_dk[c] = -1.5 / k + 2.0 * (1.0 - d * d) / (k * k * s);
    }
    double se = 0.0;
    for (int c = 0; c < C; c++) se += exp(lc[c] - lmax);
    const double lrot = -(lmax + log(se));
    const double invB = 1.0 / B;
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double gam = exp(lc[c] - lmax) / se;   // responsibility
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        // dL/dlogA_c = -gam; d (m.q) / d raw = (q - d m) / n
        for (int a = 0; a < 4; a++) {
            const double ma = r[a] / n;
            g[6 * c + a] = (float)(invB * (-gam) * dlogA_dd[c] * (q[a] - dd[c] * ma) / n);
        }
        const double sg = sigm(r[4]);
        const double dk_da = -kc[c] * 6.0 * sg * (1.0 - sg);
        g[6 * c + 4] = (float)(invB * (-gam) * dlogA_dk[c] * dk_da);
        g[6 * c + 5] = (float)(invB * (exp(logpi[c]) - gam));
    }
    // shift: -log N(t; mu, sigma^2 I) = log(2 pi) + 2 b + |t - mu|^2 / (2 sigma^2)
    const double bs = o[6 * C + 2];
    const double sig = fmin(20.0, fmax(0.2, exp(bs)));
    const bool clipped = exp(bs) < 0.2 || exp(bs) > 20.0;
    const double dx = tt[2 * bb] - o[6 * C], dy = tt[2 * bb + 1] - o[6 * C + 1];
    const double r2 = dx * dx + dy * dy;
    const double lt = log(2.0 * M_PI) + 2.0 * log(sig) + r2 / (2.0 * sig * sig);
    g[6 * C] = (float)(invB * (-dx) / (sig * sig));
    g[6 * C + 1] = (float)(invB * (-dy) / (sig * sig));
    g[6 * C + 2] = clipped ? 0.f : (float)(invB * (2.0 - r2 / (sig * sig)));
    loss[bb] = (float)(lrot + lt);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_init_params(float* __restrict__ p, size_t n, float scale, uint64_t seed,
                              uint32_t stream) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[4];
    Philox::draw(seed, (uint32_t)i, (uint32_t)(i >> 32), stream, w);
    const double u1 = Philox::u01(w[0], w[1]), u2 = Philox::u01(w[2], w[3]);
    p[i] = (float)(scale * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
}

}  // namespace thx

// ---------------------------------------------------------------------------
// Network object and C ABI
// ---------------------------------------------------------------------------
using namespace thx;

struct thx_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    int C = 8, NOUT = 51;
    size_t off[13] = {};     // parameter offsets: w1 b1 w2 b2 w3 b3 w4 b4 W5 b5 W6 b6, total
    size_t nparam = 0;
    DevBuf params, grads, adam_m, adam_v;
    DevBuf act[7];           // input, conv1..4 outputs, fc1 output, head output
    DevBuf gact[7];
    DevBuf loss, qt, tt, mix, imgs;
    DevBuf wpart;            // split-K partial weight gradients
    DevBuf wT;               // conv 1-3 weights as BF16 [co][tap][c] (proposal passes)
    long step = 0;
    int inferPrec = THX_BF16;   // operand precision of the proposal passes
};

static size_t conv_w(int l) {
    return (size_t)kConv[l].co * kConv[l].ci * kConv[l].ks * kConv[l].ks;
}

static int cnn_alloc(thx_cnn* n) {
    size_t o = 0;
    for (int l = 0; l < 4; l++) {
        n->off[2 * l] = o;
        o += conv_w(l);
        n->off[2 * l + 1] = o;
        o += kConv[l].co;
    }
    n->off[8] = o;
    o += (size_t)kFcHid * kFcIn;
    n->off[9] = o;
    o += kFcHid;
    n->off[10] = o;
    o += (size_t)n->NOUT * kFcHid;
    n->off[11] = o;
    o += n->NOUT;
    n->off[12] = o;
    n->nparam = o;
    THX_TRY(n->params.ensure(sizeof(float) * o));
    THX_TRY(n->grads.ensure(sizeof(float) * o));
    THX_TRY(n->adam_m.ensure(sizeof(float) * o));
    THX_TRY(n->adam_v.ensure(sizeof(float) * o));
    const size_t sz[7] = {(size_t)kImg * kImg, 16 * 32 * 32, 32 * 16 * 16, 64 * 8 * 8,
                          128 * 4 * 4, kFcHid, (size_t)n->NOUT};
    for (int i = 0; i < 7; i++) {
        THX_TRY(n->act[i].ensure(sizeof(float) * sz[i] * kMaxBatch));
        THX_TRY(n->gact[i].ensure(sizeof(float) * sz[i] * kMaxBatch));
    }
    THX_TRY(n->loss.ensure(sizeof(float) * kMaxBatch));
    THX_TRY(n->qt.ensure(sizeof(double) * 4 * kMaxBatch));
    THX_TRY(n->tt.ensure(sizeof(double) * 2 * kMaxBatch));
    THX_TRY(n->mix.ensure(sizeof(double) * n->NOUT * kMaxBatch));
    return THX_OK;
}

// activation buffers for forward-only passes of up to B images
static int ensure_infer(thx_cnn* n, int B) {
    const size_t sz[7] = {(size_t)kImg * kImg, 16 * 32 * 32, 32 * 16 * 16, 64 * 8 * 8,
                          128 * 4 * 4, kFcHid, (size_t)n->NOUT};
    for (int i = 1; i < 7; i++) THX_TRY(n->act[i].ensure(sizeof(float) * sz[i] * B));
    THX_TRY(n->mix.ensure(sizeof(double) * n->NOUT * B));
    return THX_OK;
}

static float* P_(thx_cnn* n, int i) { return n->params.as<float>() + n->off[i]; }
static float* G_(thx_cnn* n, int i) { return n->grads.as<float>() + n->off[i]; }

static unsigned nblk(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// Proposal passes (inference, BF16 operands, FP32 accumulation): the four
// conv layers of one image in one block, the activations kept in LDS as BF16
// with the channel innermost ([y][x][c]).  Layer 0 (1 channel in, 5x5) is a
// direct FP32 convolution over the image in LDS, each thread 4 positions x
// 16 channels from a 5 x 11 patch in registers.  A 3x3 stride-2 convolution is then a sum
// over the 9 taps of 16-channel steps: for tap (ky, kx) and channels
// 16 j + 8 h .. + 7 a lane's B fragment (output position n, lane half h) is
// one 16-byte LDS read at input pixel (2 oy - 1 + ky, 2 ox - 1 + kx), zero
// outside the image; its A fragment (output channel m) is 16 bytes of the
// weights rearranged as [co][tap][c] (k_conv_wt).  One v_mfma_f32_32x32x16_bf16
// per (32 x 32 tile, step).  The output tiles: layer 1 32 channels x 256
// positions (two per wave), layer 2 64 x 64 (one per wave), layer 3 128 x 16
// (one 32-channel tile per wave, positions padded to 32).  Bias and ReLU as
// in the FP32 layers; layer 3's output goes to global memory in the FP32
// [c][y][x] layout the fully connected layer reads.
// ---------------------------------------------------------------------------
typedef unsigned int cnn_u32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_bits(float v) {
    return __builtin_bit_cast(unsigned short, (__bf16)v);
}

// wT[(co * 9 + tap) * CI + c] = bf16(w[(co * CI + c) * 9 + tap])
__global__ void k_conv_wt(const float* __restrict__ w, int CO, int CI,
                          unsigned short* __restrict__ wT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= CO * CI * 9) return;
    const int c = i % CI, tap = (i / CI) % 9, co = i / (CI * 9);
    wT[i] = bf16_bits(w[(co * CI + c) * 9 + tap]);
}

// one (32 x 32 output tile, 16-channel step) of a 3x3 stride-2 pad-1 layer
// reading a [HI][HI][CI] BF16 input from LDS; pos = output position of this
// lane's column (-1: padding column)
template <int HI, int HO, int CI>
__device__ inline void conv_tap_step(const unsigned short* __restrict__ x,
                                     const unsigned short* __restrict__ wT, int co, int pos,
                                     int tap, int j, int h, floatx16& acc) {
    const int ky = tap / 3, kx = tap % 3;
    const cnn_u32x4 a = *reinterpret_cast<const cnn_u32x4*>(wT + (co * 9 + tap) * CI + 16 * j + 8 * h);
    cnn_u32x4 b = {0u, 0u, 0u, 0u};
    if (pos >= 0) {
        const int iy = 2 * (pos / HO) - 1 + ky, ix = 2 * (pos % HO) - 1 + kx;
        if (iy >= 0 && iy < HI && ix >= 0 && ix < HI)
            b = *reinterpret_cast<const cnn_u32x4*>(x + (iy * HI + ix) * CI + 16 * j + 8 * h);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b), acc, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_conv0123_bf16(
    const float* __restrict__ img, const float* __restrict__ w0, const float* __restrict__ b0,
    const unsigned short* __restrict__ wT1, const float* __restrict__ b1,
    const unsigned short* __restrict__ wT2, const float* __restrict__ b2,
    const unsigned short* __restrict__ wT3, const float* __restrict__ b3,
    float* __restrict__ act4) {
    using G0 = Geo<0>;
    static_assert(G0::CI == 1 && G0::KS == 5 && G0::S == 2 && G0::PD == 2 && G0::HI == 64 &&
                  G0::HO == 32 && G0::CO == 16, "layer 0 shape");
    static_assert(Geo<1>::CI == 16 && Geo<1>::HI == 32 && Geo<1>::CO == 32 && Geo<2>::CO == 64 &&
                  Geo<3>::CO == 128 && Geo<3>::HO == 4, "layer shapes");
    constexpr int W = G0::HI + 2 * G0::PD;   // 68: the zero-padded image
    // the padded image (FP32) and, once layer 0 has read it, layer 1's output
    __shared__ __attribute__((aligned(16))) float u[W * W];                    // 18.5 KB
    __shared__ __attribute__((aligned(16))) unsigned short x1[32 * 32 * 16];   // 32 KB
    static_assert(sizeof(u) >= 16 * 16 * 32 * sizeof(unsigned short), "x2 fits the image");
    float* im = u;
    unsigned short* x2 = reinterpret_cast<unsigned short*>(u);   // [16][16][32]
    unsigned short* x3 = x1;   // [8][8][64], 8 KB, once layer 1 has read x1
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, h = lane >> 5;
    // ---- layer 0 (FP32, direct): 1 -> 16 channels, 5x5 stride 2 ----
    for (int e = tid; e < W * W; e += 256) im[e] = 0.f;
    __syncthreads();
    {
        const float4* src = reinterpret_cast<const float4*>(img + (size_t)b * 4096);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;   // float4 index: row e / 16, columns 4 (e % 16)
            const float4 v = src[e];
            float* d = im + (e / 16 + G0::PD) * W + 4 * (e % 16) + G0::PD;
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    }
    __syncthreads();
    {
        const int oy = tid >> 3, ox0 = 4 * (tid & 7);
        float x[5][11];
#pragma unroll
        for (int ky = 0; ky < 5; ky++)
#pragma unroll
            for (int c = 0; c < 11; c++) x[ky][c] = im[(2 * oy + ky) * W + 2 * ox0 + c];
        // two channels at a time: 4 positions x 2 channels, stored as BF16 pairs
        for (int co = 0; co < G0::CO; co += 2) {
            float a[2][4];
#pragma unroll
            for (int d = 0; d < 2; d++) {
                a[d][0] = a[d][1] = a[d][2] = a[d][3] = b0[co + d];
#pragma unroll
                for (int ky = 0; ky < 5; ky++)
#pragma unroll
                    for (int kx = 0; kx < 5; kx++) {
                        const float wv = w0[(co + d) * 25 + ky * 5 + kx];
                        a[d][0] += wv * x[ky][kx];
                        a[d][1] += wv * x[ky][kx + 2];
                        a[d][2] += wv * x[ky][kx + 4];
                        a[d][3] += wv * x[ky][kx + 6];
                    }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                ushort2 v;
                v.x = bf16_bits(fmaxf(a[0][q], 0.f));
                v.y = bf16_bits(fmaxf(a[1][q], 0.f));
                *reinterpret_cast<ushort2*>(x1 + (oy * 32 + ox0 + q) * 16 + co) = v;
            }
        }
    }
    __syncthreads();
    // accumulator row i of a tile: channel 8 (i / 4) + 4 h + i % 4 of the tile
    // ---- layer 1: 32 channels x 256 positions, tiles 2 wave, 2 wave + 1 ----
    {
        floatx16 acc[2] = {floatx16{0}, floatx16{0}};
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int t = 0; t < 2; t++)
                conv_tap_step<32, 16, 16>(x1, wT1, l32, 32 * (2 * wave + t) + l32, tap, 0, h, acc[t]);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int pos = 32 * (2 * wave + t) + l32;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int co = 8 * g + 4 * h;
                ushort4 v;
                v.x = bf16_bits(fmaxf(acc[t][4 * g + 0] + b1[co + 0], 0.f));
                v.y = bf16_bits(fmaxf(acc[t][4 * g + 1] + b1[co + 1], 0.f));
                v.z = bf16_bits(fmaxf(acc[t][4 * g + 2] + b1[co + 2], 0.f));
                v.w = bf16_bits(fmaxf(acc[t][4 * g + 3] + b1[co + 3], 0.f));
                *reinterpret_cast<ushort4*>(x2 + pos * 32 + co) = v;
            }
        }
    }
    __syncthreads();
    // ---- layer 2: 64 channels x 64 positions, one 32 x 32 tile per wave ----
    {
        const int mt = wave & 1, nt = wave >> 1;
        floatx16 acc = {0};
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                conv_tap_step<16, 8, 32>(x2, wT2, 32 * mt + l32, 32 * nt + l32, tap, j, h, acc);
        const int pos = 32 * nt + l32;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int co = 32 * mt + 8 * g + 4 * h;
            ushort4 v;
            v.x = bf16_bits(fmaxf(acc[4 * g + 0] + b2[co + 0], 0.f));
            v.y = bf16_bits(fmaxf(acc[4 * g + 1] + b2[co + 1], 0.f));
            v.z = bf16_bits(fmaxf(acc[4 * g + 2] + b2[co + 2], 0.f));
            v.w = bf16_bits(fmaxf(acc[4 * g + 3] + b2[co + 3], 0.f));
            *reinterpret_cast<ushort4*>(x3 + pos * 64 + co) = v;
        }
    }
    __syncthreads();
    // ---- layer 3: 128 channels x 16 positions (columns 16..31 padding) ----
    {
        floatx16 acc = {0};
        const int pos = l32 < 16 ? l32 : -1;
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                conv_tap_step<8, 4, 64>(x3, wT3, 32 * wave + l32, pos, tap, j, h, acc);
        if (pos >= 0) {
            float* o = act4 + (size_t)b * 128 * 16 + pos;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int co = 32 * wave + 8 * (i >> 2) + 4 * h + (i & 3);
                o[co * 16] = fmaxf(acc[i] + b3[co], 0.f);
            }
        }
    }
}

template <int L, int WM, int WN, bool BF>
static int conv_fwd(hipStream_t st, const float* in, const float* w, const float* bias, float* out,
                    int B) {
    using G = Geo<L>;
    return gemm<WM, WN, false, RowOp, Im2col<L>, ConvStore<L>, BF>(
        st, RowOp{w, G::CI * G::KS * G::KS}, Im2col<L>{in}, ConvStore<L>{out, bias}, G::CO,
        B * G::HO * G::HO, G::CI * G::KS * G::KS);
}

// forward of B images at `in` (act[0] when null); BF (the proposal passes):
// the first layer direct in FP32, the others with BF16 operands
template <bool BF>
static int cnn_forward_p(thx_cnn* n, int B, const float* in) {
    hipStream_t st = n->stream;
    const float* x = in ? in : n->act[0].as<const float>();
    if constexpr (!BF)
        THX_TRY((conv_fwd<0, 1, 4, BF>(st, x, P_(n, 0), P_(n, 1), n->act[1].as<float>(), B)));
    if constexpr (BF) {
        // the four conv layers fused, activations in LDS (layer 1-3 weights
        // rearranged each pass: 95 k values)
        const int ci[3] = {16, 32, 64}, co[3] = {32, 64, 128};
        size_t o[4] = {0};
        for (int l = 0; l < 3; l++) o[l + 1] = o[l] + (size_t)co[l] * ci[l] * 9;
        THX_TRY(n->wT.ensure(sizeof(unsigned short) * o[3]));
        unsigned short* wT = n->wT.as<unsigned short>();
        for (int l = 0; l < 3; l++) {
            const int cnt = co[l] * ci[l] * 9;
            hipLaunchKernelGGL(k_conv_wt, dim3((cnt + 255) / 256), dim3(256), 0, st,
                               P_(n, 2 + 2 * l), co[l], ci[l], wT + o[l]);
        }
        hipLaunchKernelGGL(k_conv0123_bf16, dim3(B), dim3(256), 0, st, x, P_(n, 0), P_(n, 1),
                           wT + o[0], P_(n, 3), wT + o[1], P_(n, 5), wT + o[2], P_(n, 7),
                           n->act[4].as<float>());
        THX_HIP(hipGetLastError());
    } else {
        THX_TRY((conv_fwd<1, 1, 4, BF>(st, n->act[1].as<float>(), P_(n, 2), P_(n, 3), n->act[2].as<float>(), B)));
        THX_TRY((conv_fwd<2, 2, 2, BF>(st, n->act[2].as<float>(), P_(n, 4), P_(n, 5), n->act[3].as<float>(), B)));
        THX_TRY((conv_fwd<3, 2, 2, BF>(st, n->act[3].as<float>(), P_(n, 6), P_(n, 7), n->act[4].as<float>(), B)));
    }
    THX_TRY((gemm<2, 2, true, RowOp, RowOp, FcStore, BF>(
        st, RowOp{n->act[4].as<float>(), kFcIn}, RowOp{P_(n, 8), kFcIn},
        FcStore{n->act[5].as<float>(), P_(n, 9), kFcHid, 1}, B, kFcHid, kFcIn)));
    THX_TRY((gemm<2, 2, true, RowOp, RowOp, FcStore, BF>(
        st, RowOp{n->act[5].as<float>(), kFcHid}, RowOp{P_(n, 10), kFcHid},
        FcStore{n->act[6].as<float>(), P_(n, 11), n->NOUT, 0}, B, n->NOUT, kFcHid)));
    return THX_OK;
}

static int cnn_forward(thx_cnn* n, int B, const float* in = nullptr, bool bf16 = false) {
    return bf16 ? cnn_forward_p<true>(n, B, in) : cnn_forward_p<false>(n, B, in);
}

template <int L>
static int conv_bwd(thx_cnn* n, int B, const float* in) {
    using G = Geo<L>;
    hipStream_t st = n->stream;
    constexpr int KC = G::CI * G::KS * G::KS, HW = G::HO * G::HO;
    float* g = n->gact[L + 1].as<float>();
    const size_t no = (size_t)B * G::CO * HW;
    hipLaunchKernelGGL(k_relu_back, dim3(nblk(no)), dim3(256), 0, st, g,
                       n->act[L + 1].as<const float>(), no);
    // dW[co][k] = sum over (b, pos) of g[co] im2col[k] and db[co] = sum of g[co]
    // (the extra ones column), split over the long reduction into per-split
    // slabs, then summed in split order
    const int red = B * HW;
    const int splits = std::max(1, std::min(256, red / 1024));
    const int kPer = ((red + splits - 1) / splits + 15) / 16 * 16;
    const int nz = (red + kPer - 1) / kPer;
    const int slab = G::CO * (KC + 1);
    THX_TRY(n->wpart.ensure(sizeof(float) * (size_t)nz * slab));
    THX_TRY((gemm<1, 4, false>(st, GradRows<L>{g}, Im2colT<L>{Im2col<L>{in}},
                               WgradStore{n->wpart.as<float>(), KC + 1, slab}, G::CO, KC + 1,
                               red, splits)));
    hipLaunchKernelGGL(k_split_reduce, dim3(nblk(slab)), dim3(256), 0, st,
                       n->wpart.as<const float>(), nz, G::CO, KC, G_(n, 2 * L),
                       G_(n, 2 * L + 1));
    if (L > 0)
        THX_TRY((gemm<1, 4, false>(st, WeightT<L>{P_(n, 2 * L)}, GradCols<L>{g},
                                   GradInStore<L>{n->gact[L].as<float>()}, G::CI,
                                   B * G::HI * G::HI, G::CO * G::KS * G::KS)));
    return THX_OK;
}

// backward from gact[6] (dL/dout) to parameter gradients
static int cnn_backward(thx_cnn* n, int B, const float* in = nullptr) {
    hipStream_t st = n->stream;
    const int O = n->NOUT;
    // fc2: dW6[o][i] = sum_b g[b][o] z[b][i]; db6; dz = g W6
    THX_TRY((gemm<1, 4, false>(st, ColOp{n->gact[6].as<float>(), O}, ColOp{n->act[5].as<float>(), kFcHid},
                               PlainStore{G_(n, 10), kFcHid}, O, kFcHid, B)));
    hipLaunchKernelGGL(k_col_sum, dim3(nblk(O, 64)), dim3(64), 0, st, n->gact[6].as<float>(),
                       G_(n, 11), B, O);
    THX_TRY((gemm<2, 2, false>(st, RowOp{n->gact[6].as<float>(), O}, ColOp{P_(n, 10), kFcHid},
                               PlainStore{n->gact[5].as<float>(), kFcHid}, B, kFcHid, O)));
    hipLaunchKernelGGL(k_relu_back, dim3(nblk((size_t)B * kFcHid)), dim3(256), 0, st,
                       n->gact[5].as<float>(), n->act[5].as<const float>(), (size_t)B * kFcHid);
    // fc1
    THX_TRY((gemm<2, 2, false>(st, ColOp{n->gact[5].as<float>(), kFcHid},
                               ColOp{n->act[4].as<float>(), kFcIn}, PlainStore{G_(n, 8), kFcIn},
                               kFcHid, kFcIn, B)));
    hipLaunchKernelGGL(k_col_sum, dim3(nblk(kFcHid, 64)), dim3(64), 0, st, n->gact[5].as<float>(),
                       G_(n, 9), B, kFcHid);
    THX_TRY((gemm<2, 2, false>(st, RowOp{n->gact[5].as<float>(), kFcHid}, ColOp{P_(n, 8), kFcIn},
                               PlainStore{n->gact[4].as<float>(), kFcIn}, B, kFcIn, kFcHid)));
    const float* x0 = in ? in : n->act[0].as<const float>();
    THX_TRY(conv_bwd<3>(n, B, n->act[3].as<float>()));
    THX_TRY(conv_bwd<2>(n, B, n->act[2].as<float>()));
    THX_TRY(conv_bwd<1>(n, B, n->act[1].as<float>()));
    THX_TRY(conv_bwd<0>(n, B, x0));
    THX_HIP(hipGetLastError());
    return THX_OK;
}

namespace thx {

// Proposal parameters of every particle into mixDev, enqueued on the
// network's stream (images included) with no host synchronisation: the
// stream first waits for `ready` (the context stream's work so far), and
// `done` is recorded after the last copy for the context stream to wait on.
int cnn_propose_async(thx_cnn* n, thx_ctx* c, double* mixDev, hipEvent_t ready,
                      hipEvent_t done) {
    THX_CHECK(n && c && mixDev && c->P > 0 && c->K > 0, THX_ERR_STATE,
              "propose: particles and output needed");
    THX_CHECK(n->device == c->device, THX_ERR_ARG,
              "propose: the network and the context must be on the same device");
    THX_TRY(c->img.ensure(sizeof(float) * 4096 * (size_t)c->P));
    const int BI = std::min(kInferBatch, c->P);
    THX_TRY(ensure_infer(n, BI));
    THX_HIP(hipStreamWaitEvent(n->stream, ready, 0));
    // the images in the precision of the proposal passes
    if (n->inferPrec == THX_BF16)
        hipLaunchKernelGGL(k_images_bf16, dim3((c->P + 3) / 4), dim3(256), 0, n->stream,
                           c->dat.as<const float2>(), c->ctf.as<const float>(),
                           c->iCol.as<const int>(), c->iRow.as<const int>(), c->K, c->P,
                           c->img.as<float>());
    else
        hipLaunchKernelGGL(k_images, dim3((c->P + 3) / 4), dim3(256), 0, n->stream,
                           c->dat.as<const float2>(), c->ctf.as<const float>(),
                           c->iCol.as<const int>(), c->iRow.as<const int>(), c->K, c->P,
                           c->img.as<float>());
    THX_HIP(hipGetLastError());
    for (int b0 = 0; b0 < c->P; b0 += BI) {
        const int B = std::min(BI, c->P - b0);
        THX_TRY(cnn_forward(n, B, c->img.as<float>() + (size_t)b0 * 4096,
                            n->inferPrec == THX_BF16));
        hipLaunchKernelGGL(k_head, dim3(nblk(B, 64)), dim3(64), 0, n->stream,
                           n->act[6].as<const float>(), B, n->C, n->mix.as<double>());
        THX_HIP(hipGetLastError());
        THX_HIP(hipMemcpyAsync(mixDev + (size_t)b0 * n->NOUT, n->mix.p,
                               sizeof(double) * B * n->NOUT, hipMemcpyDeviceToDevice, n->stream));
    }
    THX_HIP(hipEventRecord(done, n->stream));
    return THX_OK;
}

}  // namespace thx

extern "C" {

int thx_cnn_create(int device, int components, uint64_t seed, thx_cnn** out) {
    clear_error();
    THX_CHECK(out && components >= 1 && components <= 16, THX_ERR_ARG, "cnn_create: bad argument");
    thx_cnn* n = new thx_cnn();
    n->device = device;
    n->C = components;
    n->NOUT = 6 * components + 3;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        set_error("cnn_create: device or stream");
        return THX_ERR_HIP;
    }
    int s = cnn_alloc(n);
    if (s != THX_OK) { delete n; return s; }
    // He-normal weights, zero biases
    for (int l = 0; l < 4; l++) {
        const size_t nw = conv_w(l);
        const float sc = sqrtf(2.f / (kConv[l].ci * kConv[l].ks * kConv[l].ks));
        hipLaunchKernelGGL(k_init_params, dim3(nblk(nw)), dim3(256), 0, n->stream, P_(n, 2 * l),
                           nw, sc, seed, (uint32_t)(100 + l));
        (void)hipMemsetAsync(P_(n, 2 * l + 1), 0, sizeof(float) * kConv[l].co, n->stream);
    }
    hipLaunchKernelGGL(k_init_params, dim3(nblk((size_t)kFcHid * kFcIn)), dim3(256), 0, n->stream,
                       P_(n, 8), (size_t)kFcHid * kFcIn, sqrtf(2.f / kFcIn), seed, 110u);
    (void)hipMemsetAsync(P_(n, 9), 0, sizeof(float) * kFcHid, n->stream);
    hipLaunchKernelGGL(k_init_params, dim3(nblk((size_t)n->NOUT * kFcHid)), dim3(256), 0, n->stream,
                       P_(n, 10), (size_t)n->NOUT * kFcHid, 0.5f * sqrtf(1.f / kFcHid), seed, 111u);
    (void)hipMemsetAsync(P_(n, 11), 0, sizeof(float) * n->NOUT, n->stream);
    (void)hipMemsetAsync(n->adam_m.p, 0, sizeof(float) * n->nparam, n->stream);
    (void)hipMemsetAsync(n->adam_v.p, 0, sizeof(float) * n->nparam, n->stream);
    if (hipStreamSynchronize(n->stream) != hipSuccess) {
        set_error("cnn_create: init kernels failed");
        return THX_ERR_HIP;
    }
    *out = n;
    return THX_OK;
}

int thx_cnn_destroy(thx_cnn* n) {
    if (!n) return THX_OK;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->stream);
    DevBuf* b[] = {&n->params, &n->grads, &n->adam_m, &n->adam_v, &n->loss, &n->qt, &n->tt,
                   &n->mix, &n->imgs, &n->wpart, &n->wT};
    for (DevBuf* x : b) x->release();
    for (int i = 0; i < 7; i++) {
        n->act[i].release();
        n->gact[i].release();
    }
    (void)hipStreamDestroy(n->stream);
    delete n;
    return THX_OK;
}

int thx_cnn_components(thx_cnn* n, int* C) {
    THX_CHECK(n && C, THX_ERR_ARG, "cnn_components: null argument");
    *C = n->C;
    return THX_OK;
}

int thx_cnn_set_precision(thx_cnn* n, int precision) {
    clear_error();
    THX_CHECK(n, THX_ERR_ARG, "cnn_set_precision: null network");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16, THX_ERR_ARG,
              "cnn_set_precision: FP32 or BF16");
    n->inferPrec = precision;
    return THX_OK;
}

int thx_cnn_num_params(thx_cnn* n, size_t* count) {
    THX_CHECK(n && count, THX_ERR_ARG, "cnn_num_params: null argument");
    *count = n->nparam;
    return THX_OK;
}

int thx_cnn_get_params(thx_cnn* n, float* p) {
    clear_error();
    THX_CHECK(n && p, THX_ERR_ARG, "cnn_get_params: null argument");
    THX_HIP(hipSetDevice(n->device));
    THX_HIP(hipStreamSynchronize(n->stream));
    THX_HIP(hipMemcpy(p, n->params.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return THX_OK;
}

int thx_cnn_set_params(thx_cnn* n, const float* p) {
    clear_error();
    THX_CHECK(n && p, THX_ERR_ARG, "cnn_set_params: null argument");
    THX_HIP(hipSetDevice(n->device));
    THX_HIP(hipStreamSynchronize(n->stream));
    THX_HIP(hipMemcpy(n->params.p, p, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    return THX_OK;
}

// Forward + loss (+ backward and parameter gradient when grad != NULL, or an
// Adam step when lr > 0).  images [B][64][64] float host, targets q [B][4],
// t [B][2] (double).  Returns the mean loss.
static int cnn_run(thx_cnn* n, const float* images, const float* dimages, int B,
                   const double* q, const double* t, double* lossOut, float* grad, float lr,
                   float* outHost) {
    THX_CHECK(n && B >= 1 && B <= kMaxBatch, THX_ERR_ARG, "cnn: batch must be in [1, 512]");
    THX_HIP(hipSetDevice(n->device));
    hipStream_t st = n->stream;
    if (images)
        THX_HIP(hipMemcpyAsync(n->act[0].p, images, sizeof(float) * B * 4096, hipMemcpyHostToDevice,
                               st));
    const float* in = images ? n->act[0].as<const float>() : dimages;
    THX_TRY(cnn_forward(n, B, in));
    if (outHost) {
        THX_HIP(hipStreamSynchronize(st));
        THX_HIP(hipMemcpy(outHost, n->act[6].p, sizeof(float) * B * n->NOUT, hipMemcpyDeviceToHost));
    }
    if (!q) return THX_OK;
    THX_HIP(hipMemcpyAsync(n->qt.p, q, sizeof(double) * 4 * B, hipMemcpyHostToDevice, st));
    THX_HIP(hipMemcpyAsync(n->tt.p, t, sizeof(double) * 2 * B, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_loss, dim3(nblk(B, 64)), dim3(64), 0, st, n->act[6].as<const float>(),
                       n->qt.as<const double>(), n->tt.as<const double>(), B, n->C,
                       n->loss.as<float>(), n->gact[6].as<float>());
    THX_HIP(hipGetLastError());
    if (grad || lr > 0.f) {
        THX_TRY(cnn_backward(n, B, in));
        if (lr > 0.f) {
            n->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
            hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparam)), dim3(256), 0, st,
                               n->params.as<float>(), n->grads.as<const float>(),
                               n->adam_m.as<float>(), n->adam_v.as<float>(), n->nparam, lr, b1, b2,
                               1e-8f, c1, c2, 1e-5f);
            THX_HIP(hipGetLastError());
        }
    }
    std::vector<float> l(B);
    THX_HIP(hipStreamSynchronize(st));
    THX_HIP(hipMemcpy(l.data(), n->loss.p, sizeof(float) * B, hipMemcpyDeviceToHost));
    if (grad) THX_HIP(hipMemcpy(grad, n->grads.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int i = 0; i < B; i++) s += l[i];
    if (lossOut) *lossOut = s / B;
    return THX_OK;
}

int thx_cnn_forward(thx_cnn* n, const float* images, int B, float* out) {
    clear_error();
    THX_CHECK(images && out, THX_ERR_ARG, "cnn_forward: null argument");
    return cnn_run(n, images, nullptr, B, nullptr, nullptr, nullptr, nullptr, 0.f, out);
}

int thx_cnn_loss_grad(thx_cnn* n, const float* images, int B, const double* q, const double* t,
                      double* loss, float* grad) {
    clear_error();
    THX_CHECK(images && q && t && loss, THX_ERR_ARG, "cnn_loss_grad: null argument");
    return cnn_run(n, images, nullptr, B, q, t, loss, grad, 0.f, nullptr);
}

int thx_cnn_train_step(thx_cnn* n, const float* images, int B, const double* q, const double* t,
                       float lr, double* loss) {
    clear_error();
    THX_CHECK(images && q && t && lr > 0.f, THX_ERR_ARG, "cnn_train_step: bad argument");
    return cnn_run(n, images, nullptr, B, q, t, loss, nullptr, lr, nullptr);
}

// Images of the context's particles (device), optional host copy [P][64][64];
// precision THX_FP32 (training, thx_ctx_images) or THX_BF16 (the transform
// the BF16 proposal passes use).
int thx_ctx_images_prec(thx_ctx* c, int precision, float* host) {
    clear_error();
    THX_CHECK(c && c->P > 0 && c->K > 0, THX_ERR_STATE, "ctx_images: no particles");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16, THX_ERR_ARG,
              "ctx_images: precision must be THX_FP32 or THX_BF16");
    THX_HIP(hipSetDevice(c->device));
    THX_TRY(c->img.ensure(sizeof(float) * 4096 * (size_t)c->P));
    if (precision == THX_BF16)
        hipLaunchKernelGGL(k_images_bf16, dim3((c->P + 3) / 4), dim3(256), 0, c->stream,
                           c->dat.as<const float2>(), c->ctf.as<const float>(),
                           c->iCol.as<const int>(), c->iRow.as<const int>(), c->K, c->P,
                           c->img.as<float>());
    else
        hipLaunchKernelGGL(k_images, dim3((c->P + 3) / 4), dim3(256), 0, c->stream,
                           c->dat.as<const float2>(), c->ctf.as<const float>(),
                           c->iCol.as<const int>(), c->iRow.as<const int>(), c->K, c->P,
                           c->img.as<float>());
    THX_HIP(hipGetLastError());
    if (host) {
        THX_HIP(hipStreamSynchronize(c->stream));
        THX_HIP(hipMemcpy(host, c->img.p, sizeof(float) * 4096 * (size_t)c->P, hipMemcpyDeviceToHost));
    }
    return THX_OK;
}

int thx_ctx_images(thx_ctx* c, float* host) { return thx_ctx_images_prec(c, THX_FP32, host); }

// Train on the context's particles (their images) with known poses: epochs
// passes of minibatches of size B in order, Adam with learning rate lr.
int thx_cnn_train_ctx(thx_cnn* n, thx_ctx* c, const double* q, const double* t, int B, int epochs,
                      float lr, double* lastLoss) {
    clear_error();
    THX_CHECK(n && c && q && t && B >= 1 && B <= kMaxBatch && epochs >= 1, THX_ERR_ARG,
              "cnn_train_ctx: bad argument");
    THX_TRY(thx_ctx_images(c, nullptr));
    THX_HIP(hipStreamSynchronize(c->stream));
    double l = 0.0;
    for (int e = 0; e < epochs; e++)
        for (int b0 = 0; b0 + B <= c->P; b0 += B)
            THX_TRY(cnn_run(n, nullptr, c->img.as<float>() + (size_t)b0 * 4096, B, q + 4 * (size_t)b0,
                            t + 2 * (size_t)b0, &l, nullptr, lr, nullptr));
    if (lastLoss) *lastLoss = l;
    return THX_OK;
}

// Proposal parameters for every particle of the context: mix[P][6C + 3].
int thx_cnn_propose_ctx(thx_cnn* n, thx_ctx* c, double* mixHost, double* mixDev) {
    clear_error();
    THX_CHECK(n && c && (mixHost || mixDev), THX_ERR_ARG, "cnn_propose_ctx: bad argument");
    // the images as the learned expectation's proposal passes make them
    THX_TRY(thx_ctx_images_prec(c, n->inferPrec == THX_BF16 ? THX_BF16 : THX_FP32, nullptr));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int BI = std::min(kInferBatch, c->P);
    THX_TRY(ensure_infer(n, BI));
    for (int b0 = 0; b0 < c->P; b0 += BI) {
        const int B = std::min(BI, c->P - b0);
        THX_TRY(cnn_forward(n, B, c->img.as<float>() + (size_t)b0 * 4096,
                            n->inferPrec == THX_BF16));
        hipLaunchKernelGGL(k_head, dim3(nblk(B, 64)), dim3(64), 0, n->stream,
                           n->act[6].as<const float>(), B, n->C, n->mix.as<double>());
        THX_HIP(hipGetLastError());
        THX_HIP(hipStreamSynchronize(n->stream));
        const size_t sz = sizeof(double) * B * n->NOUT;
        if (mixHost)
            THX_HIP(hipMemcpy(mixHost + (size_t)b0 * n->NOUT, n->mix.p, sz, hipMemcpyDeviceToHost));
        if (mixDev)
            THX_HIP(hipMemcpy(mixDev + (size_t)b0 * n->NOUT, n->mix.p, sz, hipMemcpyDeviceToDevice));
    }
    return THX_OK;
}

}  // extern "C"
