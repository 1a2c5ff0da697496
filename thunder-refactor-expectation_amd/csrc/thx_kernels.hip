// HIP kernels of the expectation step for gfx950 (CDNA4).
//
// Data layout in HBM (all row-major, pixel index fastest):
//   vol    float2 [(V/2+1) * V * V]   half-complex padded volume, x fastest
//   rotP   float2 [Rpad][Kp]          projected slices (plain path B operand)
//   rotLo  2 x 16-bit [Rpad][Kp]      low-precision B operand, shell-scaled
//   normP  float  [Rpad][Kp]          |P_rk|^2
//   S      float2 [Tpad][Kp]          translation phase factors
//   Y      float2 [Ppad][Kp]          sigRcp * ctf * X * shell scale
//   Wt     float  [Ppad][Kp]          sigRcp * ctf^2
//   Aterm  double [Ppad]              sum_k sigRcp |X|^2
//   Cterm  float  [Ppad][Rpad]        sum_k Wt |P|^2
// The log-likelihood of the reference,
//   dvp(p,r,t) = sum_k sigRcp_pk |X_pk - ctf_pk S_tk P_rk|^2
// (kernel_logDataVS, gpu/src/Kernel.cu:947-1004; logDataVSPrior_m_n_huabin,
// src/Optimiser.cpp:9666-9708) is evaluated as
//   dvp = Aterm_p + Cterm_pr - 2 Re sum_k conj(Y_pk conj(S_tk)) P_rk
// where the last sum is a dense (particle x translation) x rotation
// contraction over 2K reals that runs on MFMA: with Z = Y conj(S), the A
// operand row is (Re Z, Im Z) per pixel and the B operand (Re P, Im P), so
// sum_j A_j B_j = Re sum_k conj(Z) P.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "thx_device.hpp"
#include "thx_internal.hpp"

namespace thx {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Geometry helpers (device)
// ---------------------------------------------------------------------------

// Raw buffer resource over `bytes` bytes of device memory (gfx9 dword3:
// 32-bit data format, no swizzle).  Loads through it take 32-bit byte offsets
// and return 0 past the end instead of faulting.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             0x00020000);
}

typedef unsigned int u32x2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

// NS trilinear samples of a half-complex volume (same arithmetic as
// sample_trilinear) through a buffer resource, with every corner load of the
// NS points issued before any is used: four loads per point, one per (y, z)
// corner row, each reading the x0 and x0 + 1 voxels together (8 bytes from
// the bricked BF16 copy, LO; 16 from the FP32 volume, linear layout).  Byte
// offsets are 32-bit (volumes below 2 GiB).  Points the caller marks !ok
// read element 0 and return 0.
template <bool LO, int NS>
__device__ __forceinline__ void sample_trilinear_n(__amdgpu_buffer_rsrc_t rs, int V,
                                                   const float (&px)[NS], const float (&py)[NS],
                                                   const float (&pz)[NS], const bool (&ok)[NS],
                                                   float2 (&out)[NS]) {
    const int nc = V / 2 + 1;
    const uint32_t rowY = (uint32_t)vb_nbx(V) * 128u, rowZ = (uint32_t)vb_nby(V) * rowY;
    constexpr int NL = LO ? 2 * (3 - kVbRows) : 4;   // loads per point
    float fx[NS], fy[NS], fz[NS];
    bool cj[NS];
    uint32_t off[NS][4];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        float x = px[s], y = py[s], z = pz[s];
        cj[s] = x < 0;
        if (cj[s]) { x = -x; y = -y; z = -z; }
        if (!ok[s]) { x = 0.f; y = 0.f; z = 0.f; }
        const int x0 = (int)floorf(x), y0 = (int)floorf(y), z0 = (int)floorf(z);
        fx[s] = x - x0; fy[s] = y - y0; fz[s] = z - z0;
        const int ya = y0 < 0 ? y0 + V : y0, yb = y0 + 1 < 0 ? y0 + 1 + V : y0 + 1;
        const int za = z0 < 0 ? z0 + V : z0, zb = z0 + 1 < 0 ? z0 + 1 + V : z0 + 1;
        if constexpr (LO) {
            // byte offset of element (x, y, z): brick and in-brick parts of each
            // coordinate (all non-negative here: unsigned shifts and masks)
            constexpr uint32_t EB = 8u * kVbRows;
            const uint32_t ux = x0, uya = ya, uyb = yb, uza = za, uzb = zb;
            const uint32_t xo = (ux / kVbX) * 128u + (ux % kVbX) * EB;
            const uint32_t ya_o = (uya / kVbY) * rowY + (uya % kVbY) * (kVbX * EB);
            const uint32_t za_o = (uza / kVbZ) * rowZ + (uza % kVbZ) * (kVbX * kVbY * EB);
            const uint32_t zb_o = (uzb / kVbZ) * rowZ + (uzb % kVbZ) * (kVbX * kVbY * EB);
            if constexpr (kVbRows == 2) {
                // one element holds rows ya and ya + 1 (= yb, wrapped)
                off[s][0] = za_o + ya_o + xo;
                off[s][1] = zb_o + ya_o + xo;
            } else {
                const uint32_t yb_o = (uyb / kVbY) * rowY + (uyb % kVbY) * (kVbX * EB);
                off[s][0] = za_o + ya_o + xo;
                off[s][1] = za_o + yb_o + xo;
                off[s][2] = zb_o + ya_o + xo;
                off[s][3] = zb_o + yb_o + xo;
            }
        } else {
            off[s][0] = (uint32_t)((za * V + ya) * nc + x0) * 8u;
            off[s][1] = (uint32_t)((za * V + yb) * nc + x0) * 8u;
            off[s][2] = (uint32_t)((zb * V + ya) * nc + x0) * 8u;
            off[s][3] = (uint32_t)((zb * V + yb) * nc + x0) * 8u;
        }
    }
    // corner words per point in the reference's order: z, then y, then x;
    // each word is one BF16 complex voxel (LO) or one float (FP32, re / im)
    uint32_t w[NS][16];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if constexpr (!LO || kVbRows == 2) {
#pragma unroll
            for (int c = 0; c < NL; c++) {
                const u32x4v d = __builtin_amdgcn_raw_buffer_load_b128(rs, off[s][c], 0, 0);
                w[s][4 * c] = d.x; w[s][4 * c + 1] = d.y; w[s][4 * c + 2] = d.z; w[s][4 * c + 3] = d.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const u32x2v d = __builtin_amdgcn_raw_buffer_load_b64(rs, off[s][c], 0, 0);
                w[s][2 * c] = d.x; w[s][2 * c + 1] = d.y;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const float wz[2] = {1 - fz[s], fz[s]}, wy[2] = {1 - fy[s], fy[s]};
        const float wx[2] = {1 - fx[s], fx[s]};
        float re = 0.f, im = 0.f;
#pragma unroll
        for (int c = 0; c < 4; c++) {       // corner row: (z, y) = (c >> 1, c & 1)
            const float w0 = wx[0] * wy[c & 1] * wz[c >> 1];
            const float w1 = wx[1] * wy[c & 1] * wz[c >> 1];
            if constexpr (LO) {
                const uint32_t a = w[s][2 * c], b = w[s][2 * c + 1];
                re += w0 * __uint_as_float(a << 16);
                im += w0 * __uint_as_float(a & 0xffff0000u);
                re += w1 * __uint_as_float(b << 16);
                im += w1 * __uint_as_float(b & 0xffff0000u);
            } else {
                re += __uint_as_float(w[s][4 * c]) * w0;
                im += __uint_as_float(w[s][4 * c + 1]) * w0;
                re += __uint_as_float(w[s][4 * c + 2]) * w1;
                im += __uint_as_float(w[s][4 * c + 3]) * w1;
            }
        }
        out[s] = ok[s] ? make_float2(re, cj[s] ? -im : im) : make_float2(0.f, 0.f);
    }
}

__device__ inline unsigned short bf16_bits(float x) {
    return __builtin_bit_cast(unsigned short, (__bf16)x);
}

// Bricked BF16 copy of a half-complex volume (layout in thx_internal.hpp):
// one thread per element, round to nearest even; padding elements and the
// x + 1 voxel of the last column are 0.
__global__ void k_vol_bricked(const float2* __restrict__ vol, int V, int nbx, int nby,
                              size_t nElem, uint32_t* __restrict__ vb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nElem) return;
    const int within = (int)(e % kVbPer);
    const size_t brick = e / kVbPer;
    const int bx = (int)(brick % nbx), by = (int)((brick / nbx) % nby);
    const int bz = (int)(brick / ((size_t)nbx * nby));
    const int x = bx * kVbX + within % kVbX;
    const int y = by * kVbY + (within / kVbX) % kVbY;
    const int z = bz * kVbZ + within / (kVbX * kVbY);
    const int nc = V / 2 + 1;
    uint32_t o[2 * kVbRows] = {};
    if (x < nc && y < V && z < V) {
#pragma unroll
        for (int r = 0; r < kVbRows; r++) {
            const size_t i = ((size_t)z * V + (y + r) % V) * nc + x;
            const float2 a = vol[i];
            o[2 * r] = (uint32_t)bf16_bits(a.x) | ((uint32_t)bf16_bits(a.y) << 16);
            if (x + 1 < nc) {
                const float2 b = vol[i + 1];
                o[2 * r + 1] = (uint32_t)bf16_bits(b.x) | ((uint32_t)bf16_bits(b.y) << 16);
            }
        }
    }
#pragma unroll
    for (int w = 0; w < 2 * kVbRows; w++) vb[e * 2 * kVbRows + w] = o[w];
}

int launch_vol_bricked(hipStream_t st, const float2* vol, int V, uint32_t* vb) {
    const size_t n = vb_bytes(V) / (8 * kVbRows);
    hipLaunchKernelGGL(k_vol_bricked, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vol, V,
                       vb_nbx(V), vb_nby(V), n, vb);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ctf[p][k] from particle p's CTF attributes (CTFAttr, include/Database.h:
// 108-124: voltage, defocusU, defocusV, defocusTheta, Cs, amplitudeContrast,
// phaseShift), as allocPreCal fills ctfP with OPTIMISER_CTF_ON_THE_FLY
// (src/Optimiser.cpp:8033-8057, include/Config.h:145).  The electron
// wavelength 12.2643247 / sqrt(V (1 + 0.978466e-6 V)) Angstrom is formed in
// double and rounded to float, as the reference's RFLOAT lambda.
// Grid (P, ceil(K / 256)), particles on x.
__global__ __launch_bounds__(256) void k_ctf_attr(const float* __restrict__ attr, int K, int N,
                                                  float pixelSize, const int* __restrict__ iCol,
                                                  const int* __restrict__ iRow,
                                                  float* __restrict__ ctf) {
    const int p = blockIdx.x, k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const float* a = attr + 7 * (size_t)p;
    const float V = a[0];
    const float lambda = (float)(12.2643247 / sqrt((double)V * (1 + (double)V * 0.978466e-6)));
    ctf[(size_t)p * K + k] =
        ctf_value(iCol[k], iRow[k], pixelSize, lambda, a[1], a[2], a[3], a[4], a[5], N, a[6]);
}

int launch_ctf_attr(thx_ctx* c, const float* dAttr, int P, float pixelSize) {
    hipLaunchKernelGGL(k_ctf_attr, dim3(P, (c->K + 255) / 256), dim3(256), 0, c->stream, dAttr,
                       c->K, c->idim, pixelSize, c->iCol.as<const int>(), c->iRow.as<const int>(),
                       c->ctf.as<float>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Projection: one thread per (rotation, pixel).  The rotation matrix is
// formed in registers from the quaternion (FP64), the coordinates are cast to
// float as the reference does, and the 8 corner gathers are served from L2 /
// MALL because consecutive pixels of the set are spatial neighbours.
// Replaces kernel_getRotMat + kernel_Project3D (gpu/src/Kernel.cu:572-697).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_project(
    const float2* __restrict__ vol, int V, int pf, const double* __restrict__ quat,
    int R, const int* __restrict__ iCol, const int* __restrict__ iRow, int K,
    int Kp, float2* __restrict__ out, float* __restrict__ outNorm,
    uint32_t* __restrict__ outLo) {
    const int r = blockIdx.x;   // rotations on x, the grid dimension with the largest limit
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= Kp) return;
    const size_t o = (size_t)r * Kp + k;
    if (r >= R || k >= K) {
        out[o] = make_float2(0.f, 0.f);
        if (outNorm) outNorm[o] = 0.f;
        if (outLo) outLo[o] = 0u;
        return;
    }
    double m[3][2];
    quat_cols01(quat + 4 * (size_t)r, m);
    const double c0 = (double)(iCol[k] * pf), c1 = (double)(iRow[k] * pf);
    const float x = (float)(m[0][0] * c0 + m[0][1] * c1);
    const float y = (float)(m[1][0] * c0 + m[1][1] * c1);
    const float z = (float)(m[2][0] * c0 + m[2][1] * c1);
    const float2 v = sample_trilinear(vol, V, x, y, z);
    out[o] = v;
    if (outNorm) outNorm[o] = v.x * v.x + v.y * v.y;
    if (outLo) {   // as k_convert_lo<BF16> would convert it
        const uint32_t lo = __builtin_bit_cast(unsigned short, (__bf16)v.x);
        const uint32_t hi = __builtin_bit_cast(unsigned short, (__bf16)v.y);
        outLo[o] = lo | (hi << 16);
    }
}

int launch_project(thx_ctx* c, int cls, const double* dQuat, int R, int Rpad,
                   float2* out, float* outNorm, uint32_t* outLo) {
    if (out == c->rotP.as<float2>()) {   // the 16-bit table and the slice maxima are stale
        c->rotLoPrec = -1;
        c->shellMaxPOk = false;
    }
    THX_CHECK(!outLo || (out == c->rotP.as<float2>() && outLo == c->rotLo.as<uint32_t>() &&
                         c->rotLo.bytes >= sizeof(uint32_t) * (size_t)Rpad * c->Kp),
              THX_ERR_STATE, "project: the BF16 table goes with the context's slices");
    dim3 grid(Rpad, (c->Kp + 255) / 256);
    hipLaunchKernelGGL(k_project, grid, dim3(256), 0, c->stream,
                       c->vol[cls].as<const float2>(), c->vdim[cls], c->pf, dQuat, R,
                       c->iCol.as<const int>(), c->iRow.as<const int>(), c->K, c->Kp,
                       out, outNorm, outLo);
    THX_HIP(hipGetLastError());
    if (outLo) c->rotLoPrec = THX_BF16;
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Translation phase table (translate, src/Image/ImageFunctions.cpp:233-251;
// kernel_Translate, gpu/src/Kernel.cu:443-473): S[t][k] = exp(-i phase).
// Rows t >= T and pixels k >= K are zero so they contribute nothing.
// ---------------------------------------------------------------------------
__global__ void k_translate(const double* __restrict__ trans, int T, int Tpad,
                            const int* __restrict__ iCol, const int* __restrict__ iRow,
                            int K, int Kp, int N, float2* __restrict__ S) {
    const int t = blockIdx.x;   // translations on x, the grid dimension with the largest limit
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= Kp || t >= Tpad) return;
    float2 v = make_float2(0.f, 0.f);
    if (t < T && k < K) {
        const float rCol = (float)trans[2 * t] / N, rRow = (float)trans[2 * t + 1] / N;
        const float phase = (float)(2 * M_PI) * (iCol[k] * rCol + iRow[k] * rRow);
        v = make_float2(cosf(-phase), sinf(-phase));
    }
    S[(size_t)t * Kp + k] = v;
}

int launch_translate_raw(hipStream_t st, const double* trans, int T, int Tpad, const int* iCol,
                         const int* iRow, int K, int Kp, int N, float2* S) {
    dim3 grid(Tpad, (Kp + 255) / 256);
    hipLaunchKernelGGL(k_translate, grid, dim3(256), 0, st, trans, T, Tpad, iCol, iRow, K, Kp, N,
                       S);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// FP16 copy of the phase table (re, im packed in 32 bits) for the
// selected-pose scorer.
__global__ void k_phase_half(const float2* __restrict__ S, size_t n, uint32_t* __restrict__ S16) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = S[i];
    const unsigned short lo = __builtin_bit_cast(unsigned short, (_Float16)v.x);
    const unsigned short hi = __builtin_bit_cast(unsigned short, (_Float16)v.y);
    S16[i] = (uint32_t)lo | ((uint32_t)hi << 16);
}

int launch_phase_half(thx_ctx* c) {
    const size_t n = (size_t)c->TpadS * c->Kp;
    THX_TRY(c->S16.ensure(sizeof(uint32_t) * n));
    hipLaunchKernelGGL(k_phase_half, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->S.as<const float2>(), n, c->S16.as<uint32_t>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int launch_translate_table(thx_ctx* c) {
    dim3 grid(c->TpadS, (c->Kp + 255) / 256);
    hipLaunchKernelGGL(k_translate, grid, dim3(256), 0, c->stream, c->trans.as<const double>(),
                       c->T, c->TpadS, c->iCol.as<const int>(), c->iRow.as<const int>(), c->K,
                       c->Kp, c->idim, c->S.as<float2>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Particle operands: one block per particle.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prep(
    const float2* __restrict__ dat, const float* __restrict__ ctf,
    const float* __restrict__ sig, int P, int K, int Kp, const int* __restrict__ iSig,
    const float* __restrict__ shellScale, double* __restrict__ Aterm,
    float2* __restrict__ Y, float* __restrict__ Wt) {
    const int p = blockIdx.x;
    double a = 0.0;
    for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
        float2 y = make_float2(0.f, 0.f);
        float w = 0.f;
        if (p < P && k < K) {
            const size_t i = (size_t)p * K + k;
            const float2 x = dat[i];
            const float cc = ctf[i], s = sig[i];
            a += (double)s * ((double)x.x * x.x + (double)x.y * x.y);
            const float sc = s * cc * (shellScale ? shellScale[iSig[k]] : 1.f);
            y = make_float2(sc * x.x, sc * x.y);
            w = s * cc * cc;
        }
        Y[(size_t)p * Kp + k] = y;
        Wt[(size_t)p * Kp + k] = w;
    }
    // block reduction of a
    __shared__ double red[256];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) Aterm[p] = red[0];
}

int launch_prep_particles(thx_ctx* c, int precision) {
    THX_CHECK(c->ctfReady, THX_ERR_STATE,
              "particles have no CTF (set it with the particle arrays or thx_ctx_set_ctf)");
    THX_CHECK(c->sigReady, THX_ERR_STATE,
              "particles have no sigRcp (set it with the particle arrays or thx_ctx_init_sigma)");
    const float* scale = (precision == THX_FP16 || precision == THX_FP8)
                             ? c->shellScale.as<const float>() : nullptr;
    hipLaunchKernelGGL(k_prep, dim3(c->Ppad), dim3(256), 0, c->stream,
                       c->dat.as<const float2>(), c->ctf.as<const float>(),
                       c->sig.as<const float>(), c->P, c->K, c->Kp, c->iSig.as<const int>(),
                       scale, c->Aterm.as<double>(), c->Y.as<float2>(), c->Wt.as<float>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Per-shell scales for FP16 / FP8 operands.  The cross term is invariant to
// a per-pixel factor a_k applied as Y_k a_k and P_k / a_k; a_k is chosen per
// resolution shell so that both operands sit in the middle of the format's
// range: a = sqrt(maxP / maxY) * 2^e (e balances the two ranges).
// ---------------------------------------------------------------------------
__device__ inline void atomic_max_pos(float* addr, float v) {
    // v >= 0: the int ordering of non-negative floats matches their order
    atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
}

// shellMax[s] = max |re|, |im| over rows and the pixels of shell s: each
// thread a pixel column over every gridDim.y-th row, then per-shell maxima
// in LDS and one global atomic per touched shell and block.
constexpr int kMaxShellLds = 1024;
__global__ __launch_bounds__(256) void k_shell_max(const float2* __restrict__ src, int rows,
                                                   int Kp, int K, const int* __restrict__ iSig,
                                                   int nShell, float* __restrict__ shellMax) {
    __shared__ float sm[kMaxShellLds];
    for (int e = threadIdx.x; e < nShell; e += blockDim.x) sm[e] = 0.f;
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) {
        float m = 0.f;
#pragma unroll 8
        for (int r = blockIdx.y; r < rows; r += gridDim.y) {
            const float2 v = src[(size_t)r * Kp + k];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
        atomic_max_pos(&sm[iSig[k]], m);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nShell; e += blockDim.x)
        if (sm[e] > 0.f) atomic_max_pos(&shellMax[e], sm[e]);
}

// Per-shell power-of-two scales a_s = 2^rint(log2 sqrt(maxP_s / maxY_s)): Y is
// multiplied by a_s and the slices divided by it, so every shell's largest
// operand magnitude on both sides is sqrt(maxP_s maxY_s) and the products are
// unchanged.  FP8 (e4m3, largest finite 448) also takes one global power of
// two g that puts the largest balanced magnitude at or below 224 on both
// sides: scale[s] = a_s g, the slices get g / a_s, and the accumulator is
// multiplied by 1 / g^2.  scale[nShell] = g (1 for FP16).  One block.
__global__ void k_shell_finish(const float* __restrict__ maxP, const float* __restrict__ maxY,
                               int nShell, int fp8, float* __restrict__ scale) {
    __shared__ float red[256];
    float m = 0.f;
    for (int s = threadIdx.x; s < nShell; s += blockDim.x) {
        const float mp = maxP[s], my = maxY[s];
        float a = 1.f;
        if (mp > 0.f && my > 0.f) a = sqrtf(mp / my);
        a = exp2f(rintf(log2f(a)));   // a power of two keeps the scaling exact
        scale[s] = a;
        m = fmaxf(m, fmaxf(mp / a, my * a));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    float g = 1.f;
    if (fp8 && red[0] > 0.f) g = exp2f(floorf(log2f(224.f / red[0])));
    for (int s = threadIdx.x; s < nShell; s += blockDim.x) scale[s] *= g;
    if (threadIdx.x == 0) scale[nShell] = g;
}

// ---------------------------------------------------------------------------
// Low-precision B operand: rotLo[r][k] = P_rk / a_shell(k) in bf16 / fp16.
// ---------------------------------------------------------------------------
// FP8: two e4m3 bytes (re, im) per pixel into a 16-bit word of dst, the
// slice multiplied by g^2 / scale[s] = g / a_s.
template <int PREC>
__global__ void k_convert_lo(const float2* __restrict__ src, int rows, int Kp,
                             const int* __restrict__ iSig, int K,
                             const float* __restrict__ shellScale, int nShell,
                             uint32_t* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * Kp) return;
    const int k = (int)(i % Kp);
    float2 v = src[i];
    if (shellScale && k < K) {
        float inv = 1.f / shellScale[iSig[k]];
        if (PREC == THX_FP8) inv *= shellScale[nShell] * shellScale[nShell];
        v.x *= inv;
        v.y *= inv;
    }
    if constexpr (PREC == THX_FP8) {
        const int w = __builtin_amdgcn_cvt_pk_fp8_f32(v.x, v.y, 0, false);
        reinterpret_cast<unsigned short*>(dst)[i] = (unsigned short)(w & 0xffff);
        return;
    }
    uint32_t lo, hi;
    if (PREC == THX_BF16) {
        __bf16 a = (__bf16)v.x, b = (__bf16)v.y;
        lo = __builtin_bit_cast(unsigned short, a);
        hi = __builtin_bit_cast(unsigned short, b);
    } else {
        _Float16 a = (_Float16)v.x, b = (_Float16)v.y;
        lo = __builtin_bit_cast(unsigned short, a);
        hi = __builtin_bit_cast(unsigned short, b);
    }
    dst[i] = lo | (hi << 16);
}

// Per-shell maxima of the current slices (rotP) into shellMax[0..maxShell],
// unless already there: they depend on the projections only.
int ensure_slice_shell_max(thx_ctx* c) {
    const int nShell = c->maxShell + 1;
    THX_CHECK(nShell <= kMaxShellLds, THX_ERR_RANGE, "shell scale: too many shells");
    const void* was = c->shellMax.p;
    THX_TRY(c->shellMax.ensure(sizeof(float) * 2 * nShell));
    if (c->shellMax.p != was) c->shellMaxPOk = false;
    if (c->shellMaxPOk) return THX_OK;
    float* maxP = c->shellMax.as<float>();
    THX_HIP(hipMemsetAsync(maxP, 0, sizeof(float) * nShell, c->stream));
    // 128 row slices: each thread walks Rpad / 128 rows (16 slices left the
    // walk latency-bound: 0.14 ms for 5,000 x 3,853)
    dim3 gridP((c->K + 255) / 256, 128);
    hipLaunchKernelGGL(k_shell_max, gridP, dim3(256), 0, c->stream, c->rotP.as<const float2>(),
                       c->Rpad, c->Kp, c->K, c->iSig.as<const int>(), nShell, maxP);
    THX_HIP(hipGetLastError());
    c->shellMaxPOk = true;
    return THX_OK;
}

int launch_shell_scale(thx_ctx* c, int precision) {
    const int nShell = c->maxShell + 1;
    if (precision != THX_FP16 && precision != THX_FP8) return THX_OK;
    THX_TRY(ensure_slice_shell_max(c));
    float* maxP = c->shellMax.as<float>();
    float* maxY = maxP + nShell;
    THX_HIP(hipMemsetAsync(maxY, 0, sizeof(float) * nShell, c->stream));
    // Y with unit scale: run the prep once without scales, take its max.
    hipLaunchKernelGGL(k_prep, dim3(c->Ppad), dim3(256), 0, c->stream,
                       c->dat.as<const float2>(), c->ctf.as<const float>(),
                       c->sig.as<const float>(), c->P, c->K, c->Kp, c->iSig.as<const int>(),
                       (const float*)nullptr, c->Aterm.as<double>(), c->Y.as<float2>(),
                       c->Wt.as<float>());
    THX_HIP(hipGetLastError());
    dim3 gridY((c->K + 255) / 256, 128);   // 128 row slices, as for the slices
    hipLaunchKernelGGL(k_shell_max, gridY, dim3(256), 0, c->stream, c->Y.as<const float2>(),
                       c->Ppad, c->Kp, c->K, c->iSig.as<const int>(), nShell, maxY);
    THX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_shell_finish, dim3(1), dim3(256), 0, c->stream, maxP, maxY, nShell,
                       precision == THX_FP8 ? 1 : 0, c->shellScale.as<float>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// Per-shell bound of |P| for slices projected on the fly from a half-complex
// volume (the local scorer): a trilinear sample at pixel radius r reads
// voxels within sqrt(3) of pf r, so each voxel's magnitude bounds shells
// rint(rho / pf) - 1 .. + 1 (rho its padded radius; sqrt(3) / pf < 1 for
// pf >= 2).  Block maxima in LDS, then one atomic per shell and block.
__global__ __launch_bounds__(256) void k_vol_shell_max(const float2* __restrict__ vol, int V,
                                                       int pf, int nShell,
                                                       float* __restrict__ maxP) {
    __shared__ float sm[1024];
    for (int s = threadIdx.x; s < nShell; s += blockDim.x) sm[s] = 0.f;
    __syncthreads();
    const int nx = V / 2 + 1;
    const size_t n = (size_t)nx * V * V;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int y0 = (int)((i / nx) % V), z0 = (int)(i / ((size_t)nx * V));
        const int y = y0 < V / 2 ? y0 : y0 - V, z = z0 < V / 2 ? z0 : z0 - V;
        const int s = (int)rintf(sqrtf((float)(x * x + y * y + z * z)) / pf);
        if (s > nShell) continue;
        const float2 v = vol[i];
        const float m = fmaxf(fabsf(v.x), fabsf(v.y));
        for (int t = max(s - 1, 0); t <= min(s + 1, nShell - 1); t++)
            atomicMax(reinterpret_cast<int*>(&sm[t]), __float_as_int(m));
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nShell; s += blockDim.x)
        if (sm[s] > 0.f) atomic_max_pos(&maxP[s], sm[s]);
}

// Shell scales of the FP16 local scorer from the class volumes in use and
// the particles (as launch_shell_scale does from the grid's slices).
int launch_local_shell_scale(thx_ctx* c, const std::vector<char>& used) {
    const int nShell = c->maxShell + 1;
    THX_CHECK(nShell <= 1024, THX_ERR_RANGE, "local FP16: too many shells");
    THX_TRY(c->shellMax.ensure(sizeof(float) * 2 * nShell));
    THX_TRY(c->shellScale.ensure(sizeof(float) * (nShell + 1)));
    THX_HIP(hipMemsetAsync(c->shellMax.p, 0, sizeof(float) * 2 * nShell, c->stream));
    c->shellMaxPOk = false;   // maxP now holds volume bounds, not the slices' maxima
    float* maxP = c->shellMax.as<float>();
    float* maxY = maxP + nShell;
    for (size_t k = 0; k < used.size(); k++) {
        if (!used[k]) continue;
        hipLaunchKernelGGL(k_vol_shell_max, dim3(1024), dim3(256), 0, c->stream,
                           c->vol[k].as<const float2>(), c->vdim[k], c->pf, nShell, maxP);
        THX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_prep, dim3(c->Ppad), dim3(256), 0, c->stream,
                       c->dat.as<const float2>(), c->ctf.as<const float>(),
                       c->sig.as<const float>(), c->P, c->K, c->Kp, c->iSig.as<const int>(),
                       (const float*)nullptr, c->Aterm.as<double>(), c->Y.as<float2>(),
                       c->Wt.as<float>());
    THX_HIP(hipGetLastError());
    dim3 gridY((c->K + 255) / 256, 128);   // 128 row slices, as for the slices
    hipLaunchKernelGGL(k_shell_max, gridY, dim3(256), 0, c->stream, c->Y.as<const float2>(),
                       c->Ppad, c->Kp, c->K, c->iSig.as<const int>(), nShell, maxY);
    THX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_shell_finish, dim3(1), dim3(256), 0, c->stream, maxP, maxY, nShell, 0,
                       c->shellScale.as<float>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int launch_convert_lo(thx_ctx* c, int precision) {
    const size_t n = (size_t)c->Rpad * c->Kp;
    const float* scale = (precision == THX_FP16 || precision == THX_FP8)
                             ? c->shellScale.as<const float>() : nullptr;
    const int nShell = c->maxShell + 1;
    dim3 grid((unsigned)((n + 255) / 256));
#define THX_CONV_ARGS                                                                        \
    c->rotP.as<const float2>(), c->Rpad, c->Kp, c->iSig.as<const int>(), c->K, scale, nShell, \
        c->rotLo.as<uint32_t>()
    if (precision == THX_BF16)
        hipLaunchKernelGGL(k_convert_lo<THX_BF16>, grid, dim3(256), 0, c->stream, THX_CONV_ARGS);
    else if (precision == THX_FP16)
        hipLaunchKernelGGL(k_convert_lo<THX_FP16>, grid, dim3(256), 0, c->stream, THX_CONV_ARGS);
    else
        hipLaunchKernelGGL(k_convert_lo<THX_FP8>, grid, dim3(256), 0, c->stream, THX_CONV_ARGS);
#undef THX_CONV_ARGS
    THX_HIP(hipGetLastError());
    c->rotLoPrec = precision;
    return THX_OK;
}

// ---------------------------------------------------------------------------
// XCD-aware tile order.  Workgroups are dealt round-robin to the 8 XCDs;
// this remaps the hardware order so each XCD walks a contiguous range of
// logical tiles, which keeps one rotation block hot in that XCD's L2 while
// all particle groups stream past it.
// ---------------------------------------------------------------------------
__device__ inline int xcd_logical(int b, int nb) {
    const int q = nb / 8, rmd = nb % 8;
    const int xcd = b % 8, local = b / 8;
    return (xcd < rmd) ? xcd * (q + 1) + local : rmd * (q + 1) + (xcd - rmd) * q + local;
}

// ---------------------------------------------------------------------------
// Norm term: Cterm[p][r] = sum_k Wt[p][k] normP[r][k] (FP32 MFMA 32x32x2).
// Block tile 128 x 128, 4 waves of 64 x 64, K step 64.  blockIdx.y splits
// the pixels into ranges of kLen (a multiple of 64) when there are too few
// tiles to fill the machine; each range then writes its own slab of Cterm,
// Ppad x Rpad floats apart, summed afterwards in a fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void k_norm_gemm(
    const float* __restrict__ Wt, const float* __restrict__ normP, int Kp, int Rpad,
    int nPb, int nRb, int kLen, float* __restrict__ Cterm) {
    constexpr int LD = 65;
    __shared__ float As[128 * LD];
    __shared__ float Bs[128 * LD];
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    const int pb = L % nPb, rb = L / nPb;
    const int p0 = pb * 128, r0 = rb * 128;
    const int kBeg = blockIdx.y * kLen, kEnd = min(Kp, kBeg + kLen);
    Cterm += (size_t)blockIdx.y * nPb * 128 * Rpad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = floatx16{0};
    const int lr = tid >> 1, lh = (tid & 1) * 32;
    for (int k0 = kBeg; k0 < kEnd; k0 += 64) {
        // each thread loads 32 consecutive floats of one row of A and of B
        // (Kp and kLen are multiples of 32 and 64: a segment lies wholly
        // inside or outside the range)
        const float4* ga = reinterpret_cast<const float4*>(Wt + (size_t)(p0 + lr) * Kp + k0 + lh);
        const float4* gb = reinterpret_cast<const float4*>(normP + (size_t)(r0 + lr) * Kp + k0 + lh);
        const bool okA = (k0 + lh) < kEnd, okB = okA;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            float4 x = okA ? ga[v] : make_float4(0, 0, 0, 0);
            float4 y = okB ? gb[v] : make_float4(0, 0, 0, 0);
            float* da = As + lr * LD + lh + 4 * v;
            float* db = Bs + lr * LD + lh + 4 * v;
            da[0] = x.x; da[1] = x.y; da[2] = x.z; da[3] = x.w;
            db[0] = y.x; db[1] = y.y; db[2] = y.z; db[3] = y.w;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 64; kk += 2) {
            const int kc = kk + (lane >> 5);
            const float a0 = As[(wm + (lane & 31)) * LD + kc];
            const float a1 = As[(wm + 32 + (lane & 31)) * LD + kc];
            const float b0 = Bs[(wn + (lane & 31)) * LD + kc];
            const float b1 = Bs[(wn + 32 + (lane & 31)) * LD + kc];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = wm + 32 * mt + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int n = wn + 32 * nt + (lane & 31);
                Cterm[(size_t)(p0 + m) * Rpad + r0 + n] = acc[mt][nt][i];
            }
}

// The same norm term for the 16- and 8-bit scans on BF16 MFMA
// (v_mfma_f32_32x32x16_bf16): each FP32 operand x is split into
// hi = bf16(x) and lo = bf16(x - hi), and C = Wh.Nh + Wh.Nl + Wl.Nh, which
// leaves out Wl.Nl and the split residuals, about 2^-16 of each (non-negative)
// product: a relative error near 3e-5 of C, far below the error of the 16- or
// 8-bit cross term it is combined with.  Tiles, split-K slabs and the
// epilogue are those of k_norm_gemm; the hi and lo halves are staged in LDS
// as BF16, 8 lanes per 64-pixel row segment.
__device__ inline uint32_t bf16_hi_lo(float x, float& lo) {
    const __bf16 h = (__bf16)x;
    lo = x - (float)h;
    return __builtin_bit_cast(unsigned short, h);
}

__global__ __launch_bounds__(256, 2) void k_norm_gemm3(
    const float* __restrict__ Wt, const float* __restrict__ normP, int Kp, int Rpad,
    int nPb, int nRb, int kLen, float* __restrict__ Cterm) {
    constexpr int LD = 72;   // BF16 elements per LDS row (64 + 8: conflict-free 16-byte reads)
    __shared__ __attribute__((aligned(16))) unsigned short Ah[128 * LD], Al[128 * LD];
    __shared__ __attribute__((aligned(16))) unsigned short Bh[128 * LD], Bl[128 * LD];
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    const int pb = L % nPb, rb = L / nPb;
    const int p0 = pb * 128, r0 = rb * 128;
    const int kBeg = blockIdx.y * kLen, kEnd = min(Kp, kBeg + kLen);
    Cterm += (size_t)blockIdx.y * nPb * 128 * Rpad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hl = lane >> 5;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = floatx16{0};
    const int lr = tid >> 1, lh = (tid & 1) * 32;
    for (int k0 = kBeg; k0 < kEnd; k0 += 64) {
        // each thread splits 32 consecutive floats of one row of A and of B
        // (Kp and kLen are multiples of 32 and 64: a segment lies wholly
        // inside or outside the range)
        const bool ok = (k0 + lh) < kEnd;
        const float4* ga = reinterpret_cast<const float4*>(Wt + (size_t)(p0 + lr) * Kp + k0 + lh);
        const float4* gb = reinterpret_cast<const float4*>(normP + (size_t)(r0 + lr) * Kp + k0 + lh);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            // 8 floats of each row: one 16-byte hi store and one lo store per operand
            float x[8], y[8];
            const float4 x0 = ok ? ga[2 * v] : make_float4(0, 0, 0, 0);
            const float4 x1 = ok ? ga[2 * v + 1] : make_float4(0, 0, 0, 0);
            const float4 y0 = ok ? gb[2 * v] : make_float4(0, 0, 0, 0);
            const float4 y1 = ok ? gb[2 * v + 1] : make_float4(0, 0, 0, 0);
            x[0] = x0.x; x[1] = x0.y; x[2] = x0.z; x[3] = x0.w;
            x[4] = x1.x; x[5] = x1.y; x[6] = x1.z; x[7] = x1.w;
            y[0] = y0.x; y[1] = y0.y; y[2] = y0.z; y[3] = y0.w;
            y[4] = y1.x; y[5] = y1.y; y[6] = y1.z; y[7] = y1.w;
            u32x4 ah, al, bh, bl;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                float lo0, lo1;
                uint32_t h0 = bf16_hi_lo(x[2 * e], lo0), h1 = bf16_hi_lo(x[2 * e + 1], lo1);
                ah[e] = h0 | (h1 << 16);
                al[e] = (uint32_t)__builtin_bit_cast(unsigned short, (__bf16)lo0) |
                        ((uint32_t)__builtin_bit_cast(unsigned short, (__bf16)lo1) << 16);
                h0 = bf16_hi_lo(y[2 * e], lo0);
                h1 = bf16_hi_lo(y[2 * e + 1], lo1);
                bh[e] = h0 | (h1 << 16);
                bl[e] = (uint32_t)__builtin_bit_cast(unsigned short, (__bf16)lo0) |
                        ((uint32_t)__builtin_bit_cast(unsigned short, (__bf16)lo1) << 16);
            }
            const int off = lr * LD + lh + 8 * v;
            *reinterpret_cast<u32x4*>(Ah + off) = ah;
            *reinterpret_cast<u32x4*>(Al + off) = al;
            *reinterpret_cast<u32x4*>(Bh + off) = bh;
            *reinterpret_cast<u32x4*>(Bl + off) = bl;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 64; kk += 16) {
            const int kc = kk + 8 * hl;
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                ah[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(Ah + (wm + 32 * t + l32) * LD + kc));
                al[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(Al + (wm + 32 * t + l32) * LD + kc));
                bh[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(Bh + (wn + 32 * t + l32) * LD + kc));
                bl[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(Bl + (wn + 32 * t + l32) * LD + kc));
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mt], bh[nt], acc[mt][nt], 0, 0, 0);
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bl[nt], acc[mt][nt], 0, 0, 0);
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bh[nt], acc[mt][nt], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = wm + 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
                const int n = wn + 32 * nt + l32;
                Cterm[(size_t)(p0 + m) * Rpad + r0 + n] = acc[mt][nt][i];
            }
}

// Cterm[i] = sum over the ns slabs of part, in slab order.
__global__ void k_sum_slabs(const float* __restrict__ part, int ns, size_t n,
                            float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = part[i];
    for (int j = 1; j < ns; j++) s += part[(size_t)j * n + i];
    out[i] = s;
}

int launch_norm_gemm(thx_ctx* c, int precision) {
    const int nPb = c->Ppad / 128, nRb = c->Rpad / 128;
    // fewer tiles than two per CU: split the pixels (at least 256 per range)
    const int tiles = nPb * nRb;
    int ns = tiles >= 512 ? 1 : std::min(8, (1024 + tiles - 1) / tiles);
    ns = std::max(1, std::min(ns, c->Kp / 256));
    const int kLen = round_up((c->Kp + ns - 1) / ns, 64);
    ns = (c->Kp + kLen - 1) / kLen;
    const size_t n = (size_t)c->Ppad * c->Rpad;
    float* dst = c->Cterm.as<float>();
    if (ns > 1) {
        THX_TRY(c->normPart.ensure(sizeof(float) * n * ns));
        dst = c->normPart.as<float>();
    }
    if (precision == THX_FP32)
        hipLaunchKernelGGL(k_norm_gemm, dim3(tiles, ns), dim3(256), 0, c->stream,
                           c->Wt.as<const float>(), c->normP.as<const float>(), c->Kp, c->Rpad,
                           nPb, nRb, kLen, dst);
    else
        hipLaunchKernelGGL(k_norm_gemm3, dim3(tiles, ns), dim3(256), 0, c->stream,
                           c->Wt.as<const float>(), c->normP.as<const float>(), c->Kp, c->Rpad,
                           nPb, nRb, kLen, dst);
    THX_HIP(hipGetLastError());
    if (ns > 1) {
        hipLaunchKernelGGL(k_sum_slabs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           c->normPart.as<const float>(), ns, n, c->Cterm.as<float>());
        THX_HIP(hipGetLastError());
    }
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Scoring kernel.  A block owns PB = 128 / TC particles x TC translations
// (rows) and 128 rotations (columns).  Per K step of 32 complex pixels:
//   A[m][2k] = Re Z, A[m][2k+1] = Im Z with Z = Y[p(m)][k] conj(S[t(m)][k])
//   B[n][2k] = Re P, B[n][2k+1] = Im P
// so that sum_j A[m][j] B[n][j] = Re sum_k Z P.  Z is formed in registers
// from the particle row and the phase table, never stored in HBM.
// Epilogue: dvp = Aterm + Cterm - 2 acc, then per particle the tile maximum,
// exp(dvp - max) and the partial marginals (column sums weighted by pT,
// row sums weighted by pR) go to HBM for the FP64 merge.
// ---------------------------------------------------------------------------
template <int PREC>
struct ScoreLds {
    static constexpr bool F32 = (PREC == THX_FP32);
    static constexpr int LD = F32 ? 65 : 72;                // elements per row
    static constexpr int ESZ = F32 ? 4 : 2;
    static constexpr int AB_BYTES = 2 * 128 * LD * ESZ;
    static constexpr int D_LD = 129;
    static constexpr int D_BYTES = 128 * D_LD * 4;
    static constexpr int MAIN = AB_BYTES > D_BYTES ? AB_BYTES : D_BYTES;
};

template <int PREC>
__device__ inline unsigned short to16(float x) {
    if (PREC == THX_BF16) return __builtin_bit_cast(unsigned short, (__bf16)x);
    return __builtin_bit_cast(unsigned short, (_Float16)x);
}

template <int PREC, int TC>
__global__ __launch_bounds__(256, 2) void k_score(
    const float2* __restrict__ Y, const float2* __restrict__ S,
    const void* __restrict__ Bop, const float* __restrict__ Cterm,
    const double* __restrict__ Aterm, const float* __restrict__ pRf,
    const float* __restrict__ pTf, int P, int R, int T, int Kp, int Rpad,
    int Tpad, int nRb, int nTc, int nPg, float* __restrict__ partMax,
    float* __restrict__ partWC, int* __restrict__ partArg,
    float* __restrict__ partWR, float* __restrict__ partWT) {
    using L = ScoreLds<PREC>;
    constexpr int PB = 128 / TC;
    constexpr bool F32 = L::F32;
    constexpr int LD = L::LD;
    __shared__ __attribute__((aligned(16))) char smem[L::MAIN];
    __shared__ float redMax[PB][4];
    __shared__ int redArg[PB][4];
    __shared__ float tileMax[PB];
    __shared__ int tileArg[PB];
    __shared__ float colRed[PB][128];
    __shared__ float rowW[128];

    const int Lg = xcd_logical(blockIdx.x, gridDim.x);
    const int pg = Lg % nPg;
    const int rest = Lg / nPg;
    const int tc = rest % nTc;
    const int rb = rest / nTc;
    const int p0 = pg * PB, r0 = rb * 128, t0 = tc * TC;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = floatx16{0};

    // loader roles: row lr (0..127), half lh (16 complex pixels)
    const int lr = tid >> 1, lh = (tid & 1) * 16;
    const int lp = p0 + lr / TC, lt = t0 + lr % TC;
    const float2* yrow = Y + (size_t)lp * Kp + lh;
    const float2* srow = S + (size_t)lt * Kp + lh;
    const int lrot = r0 + lr;

    for (int k0 = 0; k0 < Kp; k0 += kChunk) {
        // ---- stage A (Z = Y conj(S)) and B into LDS ----
        if constexpr (F32) {
            float* As = reinterpret_cast<float*>(smem);
            float* Bs = As + 128 * LD;
            const float4* yv = reinterpret_cast<const float4*>(yrow + k0);
            const float4* sv = reinterpret_cast<const float4*>(srow + k0);
            const float4* bv = reinterpret_cast<const float4*>(
                static_cast<const float2*>(Bop) + (size_t)lrot * Kp + k0 + lh);
            float* da = As + lr * LD + 2 * lh;
            float* db = Bs + lr * LD + 2 * lh;
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const float4 y = yv[v], s = sv[v], b = bv[v];
                // two complex values per float4
                // Z = Y conj(S), two complex values per float4
                da[4 * v + 0] = y.x * s.x + y.y * s.y;
                da[4 * v + 1] = y.y * s.x - y.x * s.y;
                da[4 * v + 2] = y.z * s.z + y.w * s.w;
                da[4 * v + 3] = y.w * s.z - y.z * s.w;
                db[4 * v + 0] = b.x;
                db[4 * v + 1] = b.y;
                db[4 * v + 2] = b.z;
                db[4 * v + 3] = b.w;
            }
        } else {
            unsigned short* As = reinterpret_cast<unsigned short*>(smem);
            unsigned short* Bs = As + 128 * LD;
            const float4* yv = reinterpret_cast<const float4*>(yrow + k0);
            const float4* sv = reinterpret_cast<const float4*>(srow + k0);
            const uint4* bv = reinterpret_cast<const uint4*>(
                static_cast<const uint32_t*>(Bop) + (size_t)lrot * Kp + k0 + lh);
            uint32_t* da = reinterpret_cast<uint32_t*>(As + lr * LD + 2 * lh);
            uint4* db = reinterpret_cast<uint4*>(Bs + lr * LD + 2 * lh);
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const float4 y = yv[v], s = sv[v];
                const float z0r = y.x * s.x + y.y * s.y, z0i = y.y * s.x - y.x * s.y;
                const float z1r = y.z * s.z + y.w * s.w, z1i = y.w * s.z - y.z * s.w;
                da[2 * v + 0] = (uint32_t)to16<PREC>(z0r) | ((uint32_t)to16<PREC>(z0i) << 16);
                da[2 * v + 1] = (uint32_t)to16<PREC>(z1r) | ((uint32_t)to16<PREC>(z1i) << 16);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) db[v] = bv[v];
        }
        __syncthreads();
        // ---- MFMA ----
        if constexpr (F32) {
            const float* As = reinterpret_cast<const float*>(smem);
            const float* Bs = As + 128 * LD;
#pragma unroll 8
            for (int kk = 0; kk < 64; kk += 2) {
                const int kc = kk + (lane >> 5);
                const float a0 = As[(wm + (lane & 31)) * LD + kc];
                const float a1 = As[(wm + 32 + (lane & 31)) * LD + kc];
                const float b0 = Bs[(wn + (lane & 31)) * LD + kc];
                const float b1 = Bs[(wn + 32 + (lane & 31)) * LD + kc];
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
            }
        } else {
            const unsigned short* As = reinterpret_cast<const unsigned short*>(smem);
            const unsigned short* Bs = As + 128 * LD;
#pragma unroll
            for (int kk = 0; kk < 64; kk += 16) {
                const int kc = kk + 8 * (lane >> 5);
                const uint4 ra0 = *reinterpret_cast<const uint4*>(As + (wm + (lane & 31)) * LD + kc);
                const uint4 ra1 = *reinterpret_cast<const uint4*>(As + (wm + 32 + (lane & 31)) * LD + kc);
                const uint4 rb0 = *reinterpret_cast<const uint4*>(Bs + (wn + (lane & 31)) * LD + kc);
                const uint4 rb1 = *reinterpret_cast<const uint4*>(Bs + (wn + 32 + (lane & 31)) * LD + kc);
                if constexpr (PREC == THX_BF16) {
                    const bf16x8 a0 = __builtin_bit_cast(bf16x8, ra0), a1 = __builtin_bit_cast(bf16x8, ra1);
                    const bf16x8 b0 = __builtin_bit_cast(bf16x8, rb0), b1 = __builtin_bit_cast(bf16x8, rb1);
                    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
                    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
                    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
                    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
                } else {
                    const halfx8 a0 = __builtin_bit_cast(halfx8, ra0), a1 = __builtin_bit_cast(halfx8, ra1);
                    const halfx8 b0 = __builtin_bit_cast(halfx8, rb0), b1 = __builtin_bit_cast(halfx8, rb1);
                    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b0, acc[0][0], 0, 0, 0);
                    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b1, acc[0][1], 0, 0, 0);
                    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b0, acc[1][0], 0, 0, 0);
                    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b1, acc[1][1], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    // ---- epilogue: dvp tile into LDS ----
    float* Ds = reinterpret_cast<float*>(smem);
    constexpr int DLD = L::D_LD;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = wm + 32 * mt + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int n = wn + 32 * nt + (lane & 31);
                const int p = p0 + m / TC, t = t0 + m % TC, r = r0 + n;
                float v = -INFINITY;
                if (p < P && t < T && r < R)
                    v = (float)Aterm[p] + Cterm[(size_t)p * Rpad + r] - 2.f * acc[mt][nt][i];
                Ds[m * DLD + n] = v;
            }
    __syncthreads();

    // ---- per-particle maximum and its position ----
    const int cn = tid & 127, ch = tid >> 7;
#pragma unroll
    for (int pi = 0; pi < PB; pi++) {
        float mx = -INFINITY;
        int arg = -1;
        const int mb = pi * TC + ch * (TC / 2);
        for (int m = mb; m < mb + TC / 2; m++) {
            const float v = Ds[m * DLD + cn];
            if (v > mx) { mx = v; arg = m * 128 + cn; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const float o = __shfl_xor(mx, off);
            const int oa = __shfl_xor(arg, off);
            if (o > mx || (o == mx && oa < arg && oa >= 0)) { mx = o; arg = oa; }
        }
        if (lane == 0) { redMax[pi][wave] = mx; redArg[pi][wave] = arg; }
    }
    __syncthreads();
    if (tid < PB) {
        float mx = redMax[tid][0];
        int arg = redArg[tid][0];
        for (int w = 1; w < 4; w++)
            if (redMax[tid][w] > mx || (redMax[tid][w] == mx && redArg[tid][w] < arg)) {
                mx = redMax[tid][w];
                arg = redArg[tid][w];
            }
        tileMax[tid] = mx;
        tileArg[tid] = arg;
    }
    __syncthreads();

    // ---- weights and column sums (over translations, weighted by pT) ----
#pragma unroll
    for (int pi = 0; pi < PB; pi++) {
        const float mx = tileMax[pi];
        const bool live = mx > -INFINITY;
        float cs = 0.f;
        const int mb = pi * TC + ch * (TC / 2);
        for (int m = mb; m < mb + TC / 2; m++) {
            const float v = Ds[m * DLD + cn];
            const float w = live ? __expf(v - mx) : 0.f;
            Ds[m * DLD + cn] = w;
            cs += w * pTf[t0 + m % TC];
        }
        if (ch == 1) colRed[pi][cn] = cs;
        __syncthreads();
        if (ch == 0) {
            cs += colRed[pi][cn];
            const int p = p0 + pi;
            if (p < P) partWR[((size_t)p * nTc + tc) * Rpad + r0 + cn] = cs;
        }
        __syncthreads();
    }

    // ---- row sums (over rotations, weighted by pR) ----
    {
        const int m = tid >> 1, hc = (tid & 1) * 64;
        float rs = 0.f;
        for (int n = hc; n < hc + 64; n++) rs += Ds[m * DLD + n] * pRf[r0 + n];
        rs += __shfl_xor(rs, 1);
        if ((tid & 1) == 0) {
            const int p = p0 + m / TC, t = t0 + m % TC;
            if (p < P) partWT[((size_t)p * nRb + rb) * Tpad + t] = rs;
            rowW[m] = rs * pTf[t];
        }
    }
    __syncthreads();
    if (tid < PB) {
        const int p = p0 + tid;
        if (p < P) {
            float wc = 0.f;
            for (int m = tid * TC; m < (tid + 1) * TC; m++) wc += rowW[m];
            const size_t o = ((size_t)p * nTc + tc) * nRb + rb;
            partMax[o] = tileMax[tid];
            partWC[o] = wc;
            // global (r, t) of the argmax
            const int a = tileArg[tid];
            partArg[o] = a < 0 ? -1 : ((r0 + (a & 127)) * Tpad + (t0 + (a >> 7) % TC));
        }
    }
}

// ---------------------------------------------------------------------------
// 16-bit scoring kernel (BF16 / FP16 operands, FP32 accumulation).
// Block tile: MROWS rows (PB = MROWS / TC particles x TC translations) x 256
// rotations; 8 waves, each 64 rows x 256 / (8 / (MROWS / 64)) columns of
// v_mfma_f32_32x32x16 tiles (MROWS = 256: 4 x 2 waves of 64 x 128).  The
// next K step's global loads are issued before the current step's MFMAs so
// their latency hides behind matrix work; the epilogue runs on registers
// (butterfly reductions) with only small LDS exchanges.
// ---------------------------------------------------------------------------
constexpr int kNB16 = 256;
constexpr int kLD16 = 72;   // bf16 elements per LDS row (64 + 8 pad)

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// (lo, hi) -> one packed pair with a single v_cvt_pk_{bf16,f16}_f32 (RNE)
template <int PREC>
__device__ inline uint32_t pack16(float lo, float hi) {
    const floatx2 f = {lo, hi};
    if constexpr (PREC == THX_BF16)
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
    else
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, halfx2));
}

template <int PREC, int TC, int MROWS>
__global__ __launch_bounds__(512, 1) void k_score16(
    const float2* __restrict__ Y, const float2* __restrict__ S,
    const uint32_t* __restrict__ Bop, const float* __restrict__ Cterm,
    const double* __restrict__ Aterm, const float* __restrict__ pRf,
    const float* __restrict__ pTf, int P, int R, int T, int Kp, int Rpad,
    int Tpad, int nRb, int nTc, int nPg, float* __restrict__ partMax,
    float* __restrict__ partWC, int* __restrict__ partArg,
    float* __restrict__ partWR, float* __restrict__ partWT,
    const float* __restrict__ gScale) {
    constexpr int PB = MROWS / TC;
    constexpr int NB = kNB16;
    // wave grid: NWM x NWN waves, each 64 rows x WNC columns (NTW 32-column tiles)
    constexpr int NWM = MROWS / 64, NWN = 8 / NWM;
    constexpr int WNC = NB / NWN, NTW = WNC / 32;
    constexpr int NV = MROWS / 32;            // 32-row A loader groups
    constexpr bool F8 = PREC == THX_FP8;
    constexpr int LD = kLD16;                 // 16-bit elements per LDS row
    constexpr int LDB = F8 ? 80 : 2 * LD;     // bytes per LDS row (64 or 128 + 16 pad)
    constexpr int NTH = 512;
    // two LDS stages: chunk c + 1 is staged while chunk c feeds the MFMAs
    __shared__ __attribute__((aligned(16))) unsigned char AsB[2][MROWS * LDB];
    __shared__ __attribute__((aligned(16))) unsigned char BsB[2][NB * LDB];
    __shared__ float redMax[NV][NWN];     // [32-row tile][column group]
    __shared__ int redArg[NV][NWN];
    __shared__ float tileMax[PB];
    __shared__ int tileArg[PB];
    __shared__ float rowW[MROWS];
    // epilogue partials in the A stages (free once the main loop is over)
    static_assert(sizeof(float) * (NV * NB + NWN * MROWS) <= sizeof(AsB), "epilogue LDS");
    float (*colPart)[NB] = reinterpret_cast<float (*)[NB]>(&AsB[0][0]);             // [NV][NB]
    float (*rowPart)[MROWS] = reinterpret_cast<float (*)[MROWS]>(&AsB[0][0] + sizeof(float) * NV * NB);

    const int Lg = xcd_logical(blockIdx.x, gridDim.x);
    const int pg = Lg % nPg;
    const int rest = Lg / nPg;
    const int tc = rest % nTc;
    const int rb = rest / nTc;
    const int p0 = pg * PB, r0 = rb * NB, t0 = tc * TC;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wmi = wave / NWN, wni = wave % NWN;
    const int wm = wmi * 64, wn = wni * WNC;
    const int l32 = lane & 31, hl = lane >> 5;

    floatx16 acc[2][NTW];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NTW; b++) acc[a][b] = floatx16{0};

    // Loader roles (lanes of one instruction read whole contiguous rows):
    //   A: lane part = tid % 16 owns complex pixels 2 part, 2 part + 1 of the
    //      step; rows m = tid / 16 + 32 v (v = 0..3).  Row m is particle
    //      m / TC, translation m % TC: a thread needs NT = TC / 32 phase rows
    //      and 4 / NT particle rows, one float4 each.
    //      With TC = 32 (8 particles x 32 translations) a thread instead owns
    //      particles pi, pi + 4 and translations ti + 8 u (u = 0..3), with
    //      ti = rbase % 8 and pi = rbase / 8: 2 particle rows and 4 phase
    //      rows (6 loads, not 9), rows (pi + 4 j) TC + ti + 8 u.
    //   B: lane part8 = tid % 8 owns 4 complex pixels of rows tid / 8 + 64 v
    //      (FP8: lane part4 = tid % 4 owns 8 pixels of rows tid / 4 + 128 v).
    constexpr bool T32 = TC == 32;
    constexpr int NT = T32 ? 4 : TC / 32;
    constexpr int NY = T32 ? 2 : NV / NT;
    const int part = tid & 15, rbase = tid >> 4;      // rbase 0..31
    const int part8 = tid & 7, rb8 = tid >> 3;        // rb8 0..63
    const int ti = rbase & 7, pi = rbase >> 3;        // TC = 32 roles
    const float4* yp = reinterpret_cast<const float4*>(
                           Y + (size_t)(p0 + (T32 ? pi : rbase / TC)) * Kp) + part;
    const float4* sp = reinterpret_cast<const float4*>(
                           S + (size_t)(t0 + (T32 ? ti : rbase % TC)) * Kp) + part;
    const int part4 = tid & 3, rb4 = tid >> 2;        // rb4 0..127 (FP8)
    const u32x4* bp = F8 ? reinterpret_cast<const u32x4*>(
                               reinterpret_cast<const unsigned char*>(Bop) + (size_t)(r0 + rb4) * 2 * Kp) + part4
                         : reinterpret_cast<const u32x4*>(Bop + (size_t)(r0 + rb8) * Kp) + part8;
    const size_t rowF4 = (size_t)Kp / 2;
    const size_t rowU4 = F8 ? (size_t)Kp / 8 : (size_t)Kp / 4;   // B row length in 16-byte units
    // staged row of step v (v = 0 .. 7): Y row v / NT, phase row v % NT
    //   TC = 32:  particle pi + 4 (v/4), translation ti + 8 (v%4)  (NY = 2, NT = 4)
    //   TC = 64:  particle v / 2,    translation rbase + 32 (v%2) (NY = 4, NT = 2)
    //   TC = 128: particle v / 4 (= 0 .. 1), translation rbase + 32 (v%4) (NY = 2, NT = 4)

    // (plain macro, not a lambda: a lambda capturing the staging arrays by
    // reference keeps them in scratch memory instead of registers)
#define THX_GLOAD16(K0)                                                          \
    do {                                                                         \
        const int f4_ = (K0) / 2, u4_ = (K0) / 4;                                \
        _Pragma("unroll") for (int q = 0; q < NY; q++) ry[q] = yp[(size_t)(T32 ? 4 : 1) * q * rowF4 + f4_]; \
        _Pragma("unroll") for (int u = 0; u < NT; u++) rs[u] = sp[(size_t)(T32 ? 8 : 32) * u * rowF4 + f4_]; \
        if constexpr (F8) {                                                      \
            _Pragma("unroll") for (int v = 0; v < 2; v++) rbv[v] = bp[(size_t)128 * v * rowU4 + (K0) / 8]; \
        } else {                                                                 \
            _Pragma("unroll") for (int v = 0; v < 4; v++) rbv[v] = bp[(size_t)64 * v * rowU4 + u4_]; \
        }                                                                        \
    } while (0)
    float4 ry[NY], rs[NT];
    u32x4 rbv[4];
    // stage registers of one chunk into LDS buffer BUF: A = Z = Y conj(S), B as loaded
#define THX_STAGE16(BUF)                                                                       \
    do {                                                                                       \
        _Pragma("unroll") for (int v = 0; v < NV; v++) {                                       \
            const float4 y = ry[v / NT], sv = rs[v % NT];                                      \
            /* row of (Y row v / NT, phase row v % NT) */                                      \
            const int arow_ = T32 ? (pi + 4 * (v / NT)) * TC + ti + 8 * (v % NT) : rbase + 32 * v; \
            const float z0r = y.x * sv.x + y.y * sv.y, z0i = y.y * sv.x - y.x * sv.y;         \
            const float z1r = y.z * sv.z + y.w * sv.w, z1i = y.w * sv.z - y.z * sv.w;         \
            if constexpr (F8) {                                                                \
                int z8 = __builtin_amdgcn_cvt_pk_fp8_f32(z0r, z0i, 0, false);                  \
                z8 = __builtin_amdgcn_cvt_pk_fp8_f32(z1r, z1i, z8, true);                      \
                *reinterpret_cast<int*>(AsB[BUF] + arow_ * LDB + 4 * part) = z8;              \
            } else {                                                                           \
                uint2 z;                                                                       \
                z.x = pack16<PREC>(z0r, z0i);                                                  \
                z.y = pack16<PREC>(z1r, z1i);                                                  \
                *reinterpret_cast<uint2*>(AsB[BUF] + arow_ * LDB + 8 * part) = z;             \
            }                                                                                  \
        }                                                                                      \
        if constexpr (F8) {                                                                    \
            _Pragma("unroll") for (int v = 0; v < 2; v++)                                      \
                *reinterpret_cast<u32x4*>(BsB[BUF] + (rb4 + 128 * v) * LDB + 16 * part4) = rbv[v]; \
        } else {                                                                               \
            _Pragma("unroll") for (int v = 0; v < 4; v++)                                      \
                *reinterpret_cast<u32x4*>(BsB[BUF] + (rb8 + 64 * v) * LDB + 16 * part8) = rbv[v]; \
        }                                                                                      \
    } while (0)
    THX_GLOAD16(0);
    THX_STAGE16(0);
    THX_GLOAD16(kChunk);   // Kp holds at least two K steps
    __syncthreads();
    // FP8: unrolled by two (Kp is a multiple of two K steps) so that the LDS
    // buffer of each half is a constant: the MFMA half's reads of buffer buf
    // and the staging writes to buf ^ 1 are then known apart and the
    // scheduler may hoist the reads (3% faster at box 128; the 16-bit forms
    // measured slower so, and keep a run-time buffer index)
    constexpr int U = F8 ? 2 : 1;
    int bufr = 0;
    for (int k0 = 0; k0 < Kp; k0 += U * kChunk) {
#pragma unroll
    for (int h = 0; h < U; h++) {
        const int buf = U == 2 ? h : bufr;
        // buffer buf holds chunk k0 / kChunk + h.  The next chunk's staging
        // and the loads of the one after, both unconditional (past the end
        // they restage / reload the last chunk into the stage no MFMA reads):
        // conditional loads make the compiler merge old and new register
        // values with copies on every iteration
        THX_STAGE16(buf ^ 1);
        const int kn = k0 + (2 + h) * kChunk;
        THX_GLOAD16(kn < Kp ? kn : Kp - kChunk);
        // ---- MFMA on the current buffer ----
        if constexpr (F8) {
            // one 32x32x64 e4m3 step per tile: lane half hl takes bytes 32 hl .. 32 hl + 31
            // of its row in A and in B (the same k order on both sides)
            intx8 ra[2], rb8v[NTW];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                const unsigned char* src = AsB[buf] + (wm + 32 * mt + l32) * LDB + 32 * hl;
                const uint4 lo = *reinterpret_cast<const uint4*>(src);
                const uint4 hi = *reinterpret_cast<const uint4*>(src + 16);
                ra[mt] = intx8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w,
                               (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
#pragma unroll
            for (int nt = 0; nt < NTW; nt++) {
                const unsigned char* src = BsB[buf] + (wn + 32 * nt + l32) * LDB + 32 * hl;
                const uint4 lo = *reinterpret_cast<const uint4*>(src);
                const uint4 hi = *reinterpret_cast<const uint4*>(src + 16);
                rb8v[nt] = intx8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w,
                                 (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < NTW; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        ra[mt], rb8v[nt], acc[mt][nt], 0, 0, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f);
        } else {
        const unsigned short* Ab = reinterpret_cast<const unsigned short*>(AsB[buf]);
        const unsigned short* Bb = reinterpret_cast<const unsigned short*>(BsB[buf]);
#pragma unroll
        for (int kk = 0; kk < 64; kk += 16) {
            const int kc = kk + 8 * hl;
            uint4 ra[2], rb4[NTW];
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
                ra[mt] = *reinterpret_cast<const uint4*>(Ab + (wm + 32 * mt + l32) * LD + kc);
#pragma unroll
            for (int nt = 0; nt < NTW; nt++)
                rb4[nt] = *reinterpret_cast<const uint4*>(Bb + (wn + 32 * nt + l32) * LD + kc);
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < NTW; nt++) {
                    if constexpr (PREC == THX_BF16)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ra[mt]), __builtin_bit_cast(bf16x8, rb4[nt]),
                            acc[mt][nt], 0, 0, 0);
                    else
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(halfx8, ra[mt]), __builtin_bit_cast(halfx8, rb4[nt]),
                            acc[mt][nt], 0, 0, 0);
                }
        }
        }
        __syncthreads();
        if (U == 1) bufr ^= 1;
    }
    }

#undef THX_STAGE16
#undef THX_GLOAD16
    // ---- epilogue: dvp in registers ----
    // lane owns rows m = wm + 32 mt + 8 (i / 4) + 4 hl + (i % 4), columns n = wn + 32 nt + l32.
    // TC >= 32, so all rows of one 32-row tile belong to one particle.
    // FP8: undo the global operand scale g on both sides
    const float accScale = F8 ? 1.f / (gScale[0] * gScale[0]) : 1.f;
    float lmax[2];
    int larg[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const int mrow0 = wm + 32 * mt;
        const int p = p0 + mrow0 / TC;
        const bool pok = p < P;
        const float a = pok ? (float)Aterm[p] : 0.f;
        lmax[mt] = -INFINITY;
        larg[mt] = -1;
#pragma unroll
        for (int nt = 0; nt < NTW; nt++) {
            const int n = wn + 32 * nt + l32;
            const int r = r0 + n;
            const bool rok = pok && r < R;
            const float cterm = rok ? Cterm[(size_t)p * Rpad + r] : 0.f;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = mrow0 + 8 * (i >> 2) + 4 * hl + (i & 3);
                const int t = t0 + m % TC;
                float v = -INFINITY;
                if (rok && t < T) v = a + cterm - 2.f * accScale * acc[mt][nt][i];
                acc[mt][nt][i] = v;
                if (v > lmax[mt]) { lmax[mt] = v; larg[mt] = m * NB + n; }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const float o = __shfl_xor(lmax[mt], off);
            const int oa = __shfl_xor(larg[mt], off);
            if (o > lmax[mt] || (o == lmax[mt] && oa >= 0 && (larg[mt] < 0 || oa < larg[mt]))) {
                lmax[mt] = o;
                larg[mt] = oa;
            }
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            redMax[2 * wmi + mt][wni] = lmax[mt];
            redArg[2 * wmi + mt][wni] = larg[mt];
        }
    }
    __syncthreads();
    if (tid < PB) {
        constexpr int TPP = TC / 32;   // 32-row tiles per particle
        float mx = -INFINITY;
        int arg = -1;
        for (int q = tid * TPP; q < (tid + 1) * TPP; q++)
            for (int h = 0; h < NWN; h++) {
                const float v = redMax[q][h];
                const int a = redArg[q][h];
                if (v > mx || (v == mx && a >= 0 && (arg < 0 || a < arg))) { mx = v; arg = a; }
            }
        tileMax[tid] = mx;
        tileArg[tid] = arg;
    }
    __syncthreads();

    // ---- weights, column sums (x pT) and row sums (x pR) ----
    float rowv[32];   // per-lane partial row sums, index mt * 16 + i
#pragma unroll
    for (int j = 0; j < 32; j++) rowv[j] = 0.f;
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const int mrow0 = wm + 32 * mt;
        const float mx = tileMax[mrow0 / TC];
        const bool live = mx > -INFINITY;
#pragma unroll
        for (int nt = 0; nt < NTW; nt++) {
            const int n = wn + 32 * nt + l32;
            const float pr = pRf[r0 + n];
            float cs = 0.f;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int m = mrow0 + 8 * (i >> 2) + 4 * hl + (i & 3);
                const float w = live ? __expf(acc[mt][nt][i] - mx) : 0.f;
                cs += w * pTf[t0 + m % TC];
                rowv[mt * 16 + i] += w * pr;
            }
            cs += __shfl_xor(cs, 32);
            if (hl == 0) colPart[2 * wmi + mt][n] = cs;
        }
    }
    // row sums: butterfly over the 32 lanes of each half (lane l32 ends with row index l32)
#pragma unroll
    for (int s = 16; s >= 1; s >>= 1) {
        const bool up = (l32 & s) != 0;
#pragma unroll
        for (int j = 0; j < s; j++) {
            const float send = up ? rowv[j] : rowv[j + s];
            const float keep = up ? rowv[j + s] : rowv[j];
            rowv[j] = keep + __shfl_xor(send, s);
        }
    }
    {
        const int j = l32;
        const int mt = j >> 4, i = j & 15;
        const int m = wm + 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
        rowPart[wni][m] = rowv[0];
    }
    __syncthreads();
    // column sums per particle: sum of its 32-row tiles
    for (int q = tid; q < PB * NB; q += NTH) {
        const int pi = q / NB, n = q % NB;
        const int p = p0 + pi;
        float cs = 0.f;
#pragma unroll
        for (int u = 0; u < TC / 32; u++) cs += colPart[pi * (TC / 32) + u][n];
        if (p < P) partWR[((size_t)p * nTc + tc) * Rpad + r0 + n] = cs;
    }
    if (tid < MROWS) {
        const int m = tid;
        float rsum = 0.f;
#pragma unroll
        for (int h = 0; h < NWN; h++) rsum += rowPart[h][m];
        const int p = p0 + m / TC, t = t0 + m % TC;
        if (p < P) partWT[((size_t)p * nRb + rb) * Tpad + t] = rsum;
        rowW[m] = rsum * pTf[t];
    }
    __syncthreads();
    if (tid < PB) {
        const int p = p0 + tid;
        if (p < P) {
            float wc = 0.f;
            for (int m = tid * TC; m < (tid + 1) * TC; m++) wc += rowW[m];
            const size_t o = ((size_t)p * nTc + tc) * nRb + rb;
            partMax[o] = tileMax[tid];
            partWC[o] = wc;
            const int a = tileArg[tid];
            partArg[o] = a < 0 ? -1 : ((r0 + (a % NB)) * Tpad + (t0 + (a / NB) % TC));
        }
    }
}

// rows (particles x translations) per block of the 16-bit / FP8 scorer: 256,
// 8 waves of 64 x 128 (a 128-row block, 8 waves of 64 x 64, measured 1.2x
// slower for BF16 and 1.3x for FP8: every A and B fragment read from LDS now
// feeds twice the MFMAs)
constexpr int score16_rows(int) { return 256; }

template <int PREC>
static void launch_score_prec(thx_ctx* c, const void* B, int nPg, int nblk) {
#define THX_SCORE_ARGS                                                                         \
    c->Y.as<const float2>(), c->S.as<const float2>(), B, c->Cterm.as<const float>(),          \
        c->Aterm.as<const double>(), c->pRf.as<const float>(), c->pTf.as<const float>(), c->P, \
        c->R, c->T, c->Kp, c->Rpad, c->Tpad, c->nRb, c->nTc, nPg, c->partMax.as<float>(),      \
        c->partWC.as<float>(), c->partArg.as<int>(), c->partWR.as<float>(), c->partWT.as<float>()
    if constexpr (PREC == THX_FP32) {
        if (c->TC == 128)
            hipLaunchKernelGGL((k_score<PREC, 128>), dim3(nblk), dim3(256), 0, c->stream, THX_SCORE_ARGS);
        else if (c->TC == 64)
            hipLaunchKernelGGL((k_score<PREC, 64>), dim3(nblk), dim3(256), 0, c->stream, THX_SCORE_ARGS);
        else
            hipLaunchKernelGGL((k_score<PREC, 32>), dim3(nblk), dim3(256), 0, c->stream, THX_SCORE_ARGS);
    } else {
        const uint32_t* B16 = static_cast<const uint32_t*>(B);
#define THX_SCORE16_ARGS                                                                       \
    c->Y.as<const float2>(), c->S.as<const float2>(), B16, c->Cterm.as<const float>(),        \
        c->Aterm.as<const double>(), c->pRf.as<const float>(), c->pTf.as<const float>(), c->P, \
        c->R, c->T, c->Kp, c->Rpad, c->Tpad, c->nRb, c->nTc, nPg, c->partMax.as<float>(),      \
        c->partWC.as<float>(), c->partArg.as<int>(), c->partWR.as<float>(), c->partWT.as<float>(), \
        PREC == THX_FP8 ? c->shellScale.as<const float>() + c->maxShell + 1 : nullptr
        constexpr int MR = score16_rows(PREC);
        if (c->TC == 128)
            hipLaunchKernelGGL((k_score16<PREC, 128, MR>), dim3(nblk), dim3(512), 0, c->stream, THX_SCORE16_ARGS);
        else if (c->TC == 64)
            hipLaunchKernelGGL((k_score16<PREC, 64, MR>), dim3(nblk), dim3(512), 0, c->stream, THX_SCORE16_ARGS);
        else
            hipLaunchKernelGGL((k_score16<PREC, 32, MR>), dim3(nblk), dim3(512), 0, c->stream, THX_SCORE16_ARGS);
#undef THX_SCORE16_ARGS
    }
#undef THX_SCORE_ARGS
}

int score_tile_cols(int precision) { return precision == THX_FP32 ? 128 : kNB16; }

int launch_score(thx_ctx* c, int precision) {
    const int PB = (precision == THX_FP32 ? 128 : score16_rows(precision)) / c->TC;
    THX_CHECK(c->Ppad % 128 == 0 && (size_t)((c->P + PB - 1) / PB) * PB <= (size_t)c->Ppad,
              THX_ERR_STATE, "score: particle rows not padded for this tile");
    const int nPg = (c->P + PB - 1) / PB;
    const long long nblk = (long long)nPg * c->nTc * c->nRb;
    THX_CHECK(nblk < (1ll << 31), THX_ERR_RANGE, "score grid too large");
    THX_CHECK(c->NB == score_tile_cols(precision) && c->Rpad % c->NB == 0, THX_ERR_STATE,
              "score: scan shape not prepared for this precision");
    if (precision == THX_FP32)
        launch_score_prec<THX_FP32>(c, c->rotP.p, nPg, (int)nblk);
    else if (precision == THX_BF16)
        launch_score_prec<THX_BF16>(c, c->rotLo.p, nPg, (int)nblk);
    else if (precision == THX_FP16)
        launch_score_prec<THX_FP16>(c, c->rotLo.p, nPg, (int)nblk);
    else if (precision == THX_FP8)
        launch_score_prec<THX_FP8>(c, c->rotLo.p, nPg, (int)nblk);
    else {
        set_error("precision not supported by the global scorer");
        return THX_ERR_ARG;
    }
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// FP64 merge of the tile partials, one block per particle: the baseline is
// the maximum of the tile maxima; every partial is rescaled by
// exp(tileMax - baseline) (the log-sum-exp form of the streaming baseline of
// src/Optimiser.cpp:835-895 and of kernel_getMaxBase + kernel_setBaseLine +
// kernel_UpdateW, gpu/src/Kernel.cu:1052-1222).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_merge(
    const float* __restrict__ partMax, const float* __restrict__ partWC,
    const int* __restrict__ partArg, const float* __restrict__ partWR,
    const float* __restrict__ partWT, int P, int R, int T, int Rpad, int Tpad,
    int TC, int nTc, int nRb, int NB, double* __restrict__ wC, double* __restrict__ wR,
    double* __restrict__ wT, double* __restrict__ base, int* __restrict__ bestR,
    int* __restrict__ bestT, double* __restrict__ bestScore) {
    const int p = blockIdx.x;
    const int nt = nTc * nRb;
    __shared__ double f[kMaxTiles];
    __shared__ double red[256];
    __shared__ int redi[256];
    const float* pm = partMax + (size_t)p * nt;
    double mx = -INFINITY;
    int arg = -1;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
        const double v = pm[i];
        if (v > mx) { mx = v; arg = i; }
    }
    red[threadIdx.x] = mx;
    redi[threadIdx.x] = arg;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            const int oi = redi[threadIdx.x + s];
            if (o > red[threadIdx.x] || (o == red[threadIdx.x] && oi >= 0 && oi < redi[threadIdx.x])) {
                red[threadIdx.x] = o;
                redi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double M = red[0];
    const int bestTile = redi[0];
    __syncthreads();
    double wc = 0.0;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
        const double e = exp((double)pm[i] - M);
        f[i] = e;
        wc += e * partWC[(size_t)p * nt + i];
    }
    red[threadIdx.x] = wc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        wC[p] = red[0];
        base[p] = M;
        const int a = bestTile >= 0 ? partArg[(size_t)p * nt + bestTile] : -1;
        bestR[p] = a < 0 ? -1 : a / Tpad;
        bestT[p] = a < 0 ? -1 : a % Tpad;
        bestScore[p] = M;
    }
    // f index = tc * nRb + rb
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const int rb = r / NB;
        double s = 0.0;
        for (int tc = 0; tc < nTc; tc++)
            s += f[tc * nRb + rb] * partWR[((size_t)p * nTc + tc) * Rpad + r];
        wR[(size_t)p * R + r] = s;
    }
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        const int tc = t / TC;
        double s = 0.0;
        for (int rb = 0; rb < nRb; rb++)
            s += f[tc * nRb + rb] * partWT[((size_t)p * nRb + rb) * Tpad + t];
        wT[(size_t)p * T + t] = s;
    }
}

int launch_merge(thx_ctx* c, int cls) {
    ClassResult& r = c->res[cls];
    hipLaunchKernelGGL(k_merge, dim3(c->P), dim3(256), 0, c->stream, c->partMax.as<const float>(),
                       c->partWC.as<const float>(), c->partArg.as<const int>(),
                       c->partWR.as<const float>(), c->partWT.as<const float>(), c->P, c->R,
                       c->T, c->Rpad, c->Tpad, c->TC, c->nTc, c->nRb, c->NB, r.wC.as<double>(),
                       r.wR.as<double>(), r.wT.as<double>(), r.base.as<double>(),
                       r.bestR.as<int>(), r.bestT.as<int>(), r.bestScore.as<double>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Ragged local-search scorer (Optimiser::updateSP, src/Optimiser.cpp:
// 1139-1437; kernel_Project3DL + kernel_logDataVSL + kernel_UpdateWL,
// gpu/src/Kernel.cu:625-653, 827-881, 1349-1451), batched over particles.
// Each particle has its own rotations (R_p) and translations (T_p <= 128).
// A block owns one particle x up to 128 of its rotations x all its
// translations.  Per K step of 32 pixels:
//   B tile: the block's slices projected on the fly from the volume (lanes =
//           consecutive pixels of one rotation, so the 8-corner gathers of a
//           wave stay within a few cache lines);
//   A tile: Z[t][k] = Y_pk exp(+i phase_tk) = Y_pk conj(S_tk);
// then the MFMA contraction TPAD x 128 over 64 reals.  |P|^2 weighted by
// sigRcp ctf^2 accumulates in registers for the norm term.  The epilogue
// writes per-block maximum, argmax, wC, wT partials and the per-rotation wR
// (relative to the block maximum); k_merge_local combines them in FP64.
// ---------------------------------------------------------------------------
constexpr int kNBL = 128;

template <int PREC, int TPAD, int NSG>
__device__ __forceinline__ void score_local_item(int item, 
    const float2* const* __restrict__ vols, const uint32_t* const* __restrict__ volsLo,
    const int* __restrict__ pcls, int V, int pf,
    const int* __restrict__ iCol, const int* __restrict__ iRow, const int* __restrict__ iSig,
    const float* __restrict__ shellScale, int K, int Kp, int N,
    const float2* __restrict__ Y, const float* __restrict__ Wt,
    const double* __restrict__ Aterm, const double* __restrict__ quat,
    const float* __restrict__ pRf, const int* __restrict__ rOff, const int* __restrict__ nR,
    const double* __restrict__ trans, const float* __restrict__ pTf,
    const int* __restrict__ tOff, const int* __restrict__ nT, const int2* __restrict__ blocks,
    float* __restrict__ bMax, float* __restrict__ bWC, int* __restrict__ bArg,
    float* __restrict__ bWT, float* __restrict__ wRrel) {
    constexpr bool F32 = (PREC == THX_FP32);
    constexpr int LD = F32 ? 65 : 72;
    using E = typename std::conditional<F32, float, unsigned short>::type;
    constexpr int MT = TPAD / 32;                   // 32-row MFMA tiles
    __shared__ __attribute__((aligned(16))) E As[TPAD * LD];
    __shared__ __attribute__((aligned(16))) E Bs[kNBL * LD];
    __shared__ float trs[TPAD][2];
    __shared__ float Cs[kNBL];
    __shared__ float Cpart[4][kNBL];
    __shared__ float redMax[4];
    __shared__ int redArg[4];
    __shared__ float rowPart[4][TPAD];

    const int2 bi = blocks[item];
    const int p = bi.x, r0 = bi.y;
    const int Rp = nR[p], Tp = nT[p];
    const int nr = min(kNBL, Rp - r0);
    const int rbase = rOff[p] + r0, tbase = tOff[p];
    const int cl = pcls ? pcls[p] : 0;
    // the class volume (FP32) or its BF16 copy as a buffer resource
    const uint32_t vbytes = F32 ? (uint32_t)(V / 2 + 1) * V * V * 8u : (uint32_t)vb_bytes(V);
    const __amdgpu_buffer_rsrc_t vrs =
        F32 ? buf_rsrc(vols[cl], vbytes) : buf_rsrc(volsLo[cl], vbytes);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hl = lane >> 5;

    // each lane owns two of the item's rotations, lane and lane + 64, with
    // their first two matrix columns in registers (zero past nr)
    double rm[2][6];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        double m[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        if (lane + 64 * h < nr) quat_cols01(quat + 4 * (size_t)(rbase + lane + 64 * h), m);
        rm[h][0] = m[0][0]; rm[h][1] = m[0][1];
        rm[h][2] = m[1][0]; rm[h][3] = m[1][1];
        rm[h][4] = m[2][0]; rm[h][5] = m[2][1];
    }
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    if (tid < TPAD) {
        float tx = 0.f, ty = 0.f;
        if (tid < Tp) {
            tx = (float)trans[2 * (size_t)(tbase + tid)] / N;
            ty = (float)trans[2 * (size_t)(tbase + tid) + 1] / N;
        }
        trs[tid][0] = tx;
        trs[tid][1] = ty;
    }
    __syncthreads();

    floatx16 acc[MT];
#pragma unroll
    for (int a = 0; a < MT; a++) acc[a] = floatx16{0};
    float cacc[2] = {0.f, 0.f};   // norm-term sums of the lane's two rotations
    const float2* yrow = Y + (size_t)p * Kp;
    const float* wrow = Wt + (size_t)p * Kp;
    const int grp = tid >> 5;            // 8 groups of 32 lanes

    for (int k0 = 0; k0 < Kp; k0 += kChunk) {
        const int k = k0 + l32;
        const int ic = iCol[k], ir = iRow[k];
        // ---- A: Z = Y conj(S) ----
        {
            const float2 y = yrow[k];
#pragma unroll
            for (int j = 0; j < TPAD / 8; j++) {
                const int t = grp + 8 * j;
                const float ph = (float)(2 * M_PI) * (ic * trs[t][0] + ir * trs[t][1]);
                float sn, cs;
                sincosf(ph, &sn, &cs);
                const float zr = y.x * cs - y.y * sn, zi = y.x * sn + y.y * cs;
                if constexpr (F32) {
                    As[t * LD + 2 * l32] = zr;
                    As[t * LD + 2 * l32 + 1] = zi;
                } else {
                    *reinterpret_cast<uint32_t*>(As + t * LD + 2 * l32) = pack16<PREC>(zr, zi);
                }
            }
        }
        // ---- B: slices of the item's rotations at the chunk's pixels ----
        // wave w fills pixels 8w .. 8w + 7 of the chunk for all 128 rotations:
        // the 64 lanes of one load gather the same pixel of 64 neighbouring
        // rotations (the item's rotations are a compact cluster in pose
        // order), so at the low and middle radii they share cache lines
        {
            const int kb = k0 + 8 * wv;
            // NS samples per step with all their corner loads in flight (NSG for
            // the 16-bit forms, a launch-shape choice of the tuner)
            constexpr int NS = F32 ? 2 : NSG;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int rr = lane + 64 * h;
                const bool rok = rr < nr;
                float2 v8[8];
#pragma unroll
                for (int j0 = 0; j0 < 8; j0 += NS) {
                    float px[NS], py[NS], pz[NS];
                    bool ok[NS];
                    float2 v[NS];
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        const int kk = kb + j0 + s;
                        const double c0 = (double)(iCol[kk] * pf), c1 = (double)(iRow[kk] * pf);
                        ok[s] = rok && kk < K;
                        px[s] = (float)(rm[h][0] * c0 + rm[h][1] * c1);
                        py[s] = (float)(rm[h][2] * c0 + rm[h][3] * c1);
                        pz[s] = (float)(rm[h][4] * c0 + rm[h][5] * c1);
                    }
                    sample_trilinear_n<!F32, NS>(vrs, V, px, py, pz, ok, v);
#pragma unroll
                    for (int s = 0; s < NS; s++) v8[j0 + s] = v[s];
                }
#pragma unroll
                for (int j = 0; j < 8; j++)
                    cacc[h] += wrow[kb + j] * (v8[j].x * v8[j].x + v8[j].y * v8[j].y);
                if constexpr (F32) {
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        Bs[rr * LD + 2 * (8 * wv + j)] = v8[j].x;
                        Bs[rr * LD + 2 * (8 * wv + j) + 1] = v8[j].y;
                    }
                } else {
                    // FP16: the slice is divided by its shell's scale (Y carries it)
                    uint32_t pk[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        float inv = 1.f;
                        if constexpr (PREC == THX_FP16) inv = 1.f / shellScale[iSig[kb + j]];
                        pk[j] = pack16<PREC>(v8[j].x * inv, v8[j].y * inv);
                    }
                    uint4* dst = reinterpret_cast<uint4*>(Bs + rr * LD + 16 * wv);
                    dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
                    dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
                }
            }
        }
        __syncthreads();
        // ---- MFMA: wave owns columns 32 wave .. +32 ----
        const int nrow = 32 * wave + l32;
        if constexpr (F32) {
#pragma unroll 4
            for (int kk = 0; kk < 64; kk += 2) {
                const float b = Bs[nrow * LD + kk + hl];
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    const float a = As[(32 * mt + l32) * LD + kk + hl];
                    acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[mt], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 64; kk += 16) {
                const int kc = kk + 8 * hl;
                const uint4 rb4 = *reinterpret_cast<const uint4*>(Bs + nrow * LD + kc);
#pragma unroll
                for (int mt = 0; mt < MT; mt++) {
                    const uint4 ra = *reinterpret_cast<const uint4*>(As + (32 * mt + l32) * LD + kc);
                    if constexpr (PREC == THX_BF16)
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ra), __builtin_bit_cast(bf16x8, rb4), acc[mt], 0, 0, 0);
                    else
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(halfx8, ra), __builtin_bit_cast(halfx8, rb4), acc[mt], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    // ---- norm term per rotation: the four waves' pixel sums ----
    Cpart[wave][lane] = cacc[0];
    Cpart[wave][lane + 64] = cacc[1];
    __syncthreads();
    if (tid < kNBL) Cs[tid] = (Cpart[0][tid] + Cpart[1][tid]) + (Cpart[2][tid] + Cpart[3][tid]);
    __syncthreads();

    // ---- dvp, block maximum ----
    const float a = (float)Aterm[p];
    const int ncol = 32 * wave + l32;
    const bool cok = ncol < nr;
    const float cterm = Cs[ncol];
    float lmax = -INFINITY;
    int larg = -1;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            float v = -INFINITY;
            if (cok && t < Tp) v = a + cterm - 2.f * acc[mt][i];
            acc[mt][i] = v;
            if (v > lmax) { lmax = v; larg = ncol * TPAD + t; }
        }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(lmax, off);
        const int oa = __shfl_xor(larg, off);
        if (o > lmax || (o == lmax && oa >= 0 && (larg < 0 || oa < larg))) { lmax = o; larg = oa; }
    }
    if (lane == 0) { redMax[wave] = lmax; redArg[wave] = larg; }
    __syncthreads();
    float M = redMax[0];
    int arg = redArg[0];
    for (int w = 1; w < 4; w++)
        if (redMax[w] > M || (redMax[w] == M && redArg[w] >= 0 && (arg < 0 || redArg[w] < arg))) {
            M = redMax[w];
            arg = redArg[w];
        }
    const bool live = M > -INFINITY;

    // ---- weights, wR (column sums x pT), wT (row sums x pR) ----
    constexpr int NV = MT * 16 < 32 ? 32 : MT * 16;   // row values per lane (padded)
    float rowv[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) rowv[j] = 0.f;
    const float pr = cok ? pRf[rbase + ncol] : 0.f;
    float cs = 0.f;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            const float w = live ? __expf(acc[mt][i] - M) : 0.f;
            cs += w * (t < Tp ? pTf[tbase + t] : 0.f);
            rowv[mt * 16 + i] = w * pr;
        }
    cs += __shfl_xor(cs, 32);
    if (hl == 0 && cok) wRrel[rbase + ncol] = cs;
    // butterfly: groups of 32 values; lane l32 ends with value index g*32 + l32 of each group
    constexpr int NG = NV / 32;
#pragma unroll
    for (int g = 0; g < NG; g++) {
#pragma unroll
        for (int s = 16; s >= 1; s >>= 1) {
            const bool up = (l32 & s) != 0;
#pragma unroll
            for (int j = 0; j < s; j++) {
                const float send = up ? rowv[g * 32 + j] : rowv[g * 32 + j + s];
                const float keep = up ? rowv[g * 32 + j + s] : rowv[g * 32 + j];
                rowv[g * 32 + j] = keep + __shfl_xor(send, s);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int jv = g * 32 + l32;          // value index = mt * 16 + i
        const int mt = jv >> 4, i = jv & 15;
        if (mt < MT) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            rowPart[wave][t] = rowv[g * 32];
        }
    }
    __syncthreads();
    __shared__ float rowW[TPAD];
    if (tid < TPAD) {
        const float rsum = rowPart[0][tid] + rowPart[1][tid] + rowPart[2][tid] + rowPart[3][tid];
        bWT[(size_t)item * kLocalWtStride + tid] = rsum;
        rowW[tid] = tid < Tp ? rsum * pTf[tbase + tid] : 0.f;
    }
    __syncthreads();
    if (tid == 0) {
        float wc = 0.f;
        for (int t = 0; t < TPAD; t++) wc += rowW[t];
        bMax[item] = M;
        bWC[item] = wc;
        bArg[item] = arg < 0 ? -1 : ((r0 + arg / TPAD) * kNBL + arg % TPAD);
    }
}


// Direct launch: one block per work item (particle, 128 rotations).
template <int PREC, int TPAD, int NSG>
__global__ __launch_bounds__(256, 2) void k_score_local(
    const float2* const* __restrict__ vols, const uint32_t* const* __restrict__ volsLo,
    const int* __restrict__ pcls, int V, int pf,
    const int* __restrict__ iCol, const int* __restrict__ iRow, const int* __restrict__ iSig,
    const float* __restrict__ shellScale, int K, int Kp, int N,
    const float2* __restrict__ Y, const float* __restrict__ Wt,
    const double* __restrict__ Aterm, const double* __restrict__ quat,
    const float* __restrict__ pRf, const int* __restrict__ rOff, const int* __restrict__ nR,
    const double* __restrict__ trans, const float* __restrict__ pTf,
    const int* __restrict__ tOff, const int* __restrict__ nT, const int2* __restrict__ blocks,
    float* __restrict__ bMax, float* __restrict__ bWC, int* __restrict__ bArg,
    float* __restrict__ bWT, float* __restrict__ wRrel, const int* __restrict__ nDev) {
    // nDev: the item count in device memory (a list built on the device; the
    // grid is its upper bound), else one block per item
    const int n = nDev ? *nDev : (int)gridDim.x;
    if ((int)blockIdx.x >= n) return;
    // the list is in pose order: each XCD walks a contiguous range of it, so
    // concurrent blocks of one XCD gather from overlapping volume regions
    score_local_item<PREC, TPAD, NSG>(xcd_logical(blockIdx.x, n), vols, volsLo, pcls, V, pf, iCol, iRow, iSig, shellScale, K, Kp, N, Y, Wt, Aterm, quat, pRf, rOff, nR, trans, pTf, tOff, nT, blocks, bMax, bWC, bArg, bWT, wRrel);
}

// Persistent launch: a grid sized to the resident capacity pulls work items
// from atomic counters until the ragged list is exhausted.  The list (in pose
// order) is cut into 8 contiguous ranges, one per XCD (workgroups are dealt
// to the XCDs round-robin, so block b starts on range b % 8); a block that
// drains its range moves on to the next ones.
template <int PREC, int TPAD, int NSG>
__global__ __launch_bounds__(256, 2) void k_score_local_q(
    const float2* const* __restrict__ vols, const uint32_t* const* __restrict__ volsLo,
    const int* __restrict__ pcls, int V, int pf,
    const int* __restrict__ iCol, const int* __restrict__ iRow, const int* __restrict__ iSig,
    const float* __restrict__ shellScale, int K, int Kp, int N,
    const float2* __restrict__ Y, const float* __restrict__ Wt,
    const double* __restrict__ Aterm, const double* __restrict__ quat,
    const float* __restrict__ pRf, const int* __restrict__ rOff, const int* __restrict__ nR,
    const double* __restrict__ trans, const float* __restrict__ pTf,
    const int* __restrict__ tOff, const int* __restrict__ nT, const int2* __restrict__ blocks,
    float* __restrict__ bMax, float* __restrict__ bWC, int* __restrict__ bArg,
    float* __restrict__ bWT, float* __restrict__ wRrel, int nItemsArg,
    int* __restrict__ counter, const int* __restrict__ nDev) {
    __shared__ int item;
    const int nItems = nDev ? *nDev : nItemsArg;   // nDev: count in device memory
    const int q = nItems / 8, rmd = nItems % 8;
    int h = 0, x = blockIdx.x % 8;
    for (;;) {
        const int lo = x * q + min(x, rmd), n = q + (x < rmd ? 1 : 0);
        __syncthreads();   // the previous item's LDS traffic is over
        if (threadIdx.x == 0) item = atomicAdd(&counter[x], 1);
        __syncthreads();
        const int j = __builtin_amdgcn_readfirstlane(item);   // block-uniform: scalar registers
        if (j >= n) {      // this range is drained: the next XCD's
            if (++h == 8) break;
            x = (x + 1) % 8;
            continue;
        }
        const int it = lo + j;
        score_local_item<PREC, TPAD, NSG>(it, vols, volsLo, pcls, V, pf, iCol, iRow, iSig, shellScale, K, Kp, N, Y, Wt, Aterm, quat, pRf, rOff, nR, trans, pTf, tOff, nT, blocks, bMax, bWC, bArg, bWT, wRrel);
    }
}

// counter[0..7] = 0 (a kernel, so a captured graph holds no memset node)
__global__ void k_zero_int(int* __restrict__ counter) {
    if (threadIdx.x < 8 && blockIdx.x == 0) counter[threadIdx.x] = 0;
}

template <int PREC, int NSG>
static void launch_local_prec(hipStream_t st, int TPAD, int nblk, int mode, int grid, int* counter,
                              const float2* const* vols, const uint32_t* const* volsLo,
                              const int* pcls, int V, int pf,
                              const int* iCol, const int* iRow, const int* iSig,
                              const float* shellScale, int K, int Kp, int N,
                              const float2* Y, const float* Wt, const double* Aterm,
                              const double* quat, const float* pRf, const int* rOff,
                              const int* nR, const double* trans, const float* pTf,
                              const int* tOff, const int* nT, const int2* blocks, float* bMax,
                              float* bWC, int* bArg, float* bWT, float* wRrel, const int* nDev) {
#define THX_LOCAL_ARGS vols, volsLo, pcls, V, pf, iCol, iRow, iSig, shellScale, K, Kp, N, Y, Wt, Aterm, \
                       quat, pRf, rOff, \
                       nR, trans, pTf, tOff, nT, blocks, bMax, bWC, bArg, bWT, wRrel
    if (mode == kLocalQueue) {
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, counter);
        const int g = std::min(grid, nblk);
        if (TPAD == 32)
            hipLaunchKernelGGL((k_score_local_q<PREC, 32, NSG>), dim3(g), dim3(256), 0, st, THX_LOCAL_ARGS, nblk, counter, nDev);
        else if (TPAD == 64)
            hipLaunchKernelGGL((k_score_local_q<PREC, 64, NSG>), dim3(g), dim3(256), 0, st, THX_LOCAL_ARGS, nblk, counter, nDev);
        else
            hipLaunchKernelGGL((k_score_local_q<PREC, 128, NSG>), dim3(g), dim3(256), 0, st, THX_LOCAL_ARGS, nblk, counter, nDev);
    } else {
        if (TPAD == 32)
            hipLaunchKernelGGL((k_score_local<PREC, 32, NSG>), dim3(nblk), dim3(256), 0, st, THX_LOCAL_ARGS, nDev);
        else if (TPAD == 64)
            hipLaunchKernelGGL((k_score_local<PREC, 64, NSG>), dim3(nblk), dim3(256), 0, st, THX_LOCAL_ARGS, nDev);
        else
            hipLaunchKernelGGL((k_score_local<PREC, 128, NSG>), dim3(nblk), dim3(256), 0, st, THX_LOCAL_ARGS, nDev);
    }
#undef THX_LOCAL_ARGS
}

// resident work-item capacity of the persistent launch: blocks per CU x CUs
template <int PREC, int NSG>
static const void* local_queue_fn(int TPAD) {
    return TPAD == 32 ? (const void*)k_score_local_q<PREC, 32, NSG>
           : TPAD == 64 ? (const void*)k_score_local_q<PREC, 64, NSG>
                        : (const void*)k_score_local_q<PREC, 128, NSG>;
}

static int local_queue_grid(int precision, int TPAD, int ns) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1024;
    const void* fn = precision == THX_FP32 ? local_queue_fn<THX_FP32, 4>(TPAD)
                     : precision == THX_BF16 ? (ns == 8 ? local_queue_fn<THX_BF16, 8>(TPAD)
                                                        : local_queue_fn<THX_BF16, 4>(TPAD))
                                             : (ns == 8 ? local_queue_fn<THX_FP16, 8>(TPAD)
                                                        : local_queue_fn<THX_FP16, 4>(TPAD));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0) != hipSuccess || per < 1)
        per = 2;
    return cus * per;
}

int launch_score_local(thx_ctx* c, int precision, int TPAD, int nblk, const LocalBatch& b,
                       int shape, const int* nItemsDev) {
    const int mode = local_shape_mode(shape), ns = local_shape_ns(shape);
    THX_CHECK(mode == kLocalDirect || mode == kLocalQueue, THX_ERR_ARG,
              "local scorer: unknown launch shape");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16 || precision == THX_FP16,
              THX_ERR_ARG, "local scorer supports FP32, BF16 and FP16");
    THX_CHECK(precision == THX_FP32 || b.volsLo, THX_ERR_STATE,
              "local scorer: the 16-bit forms need the BF16 volume copies");
    {
        // the gathers address the volume with 32-bit byte offsets
        const long long Vd = (long long)c->idim * c->pf;
        THX_CHECK((Vd / 2 + 1) * Vd * Vd * 8 < (1ll << 31) && vb_bytes((int)Vd) < (1ull << 31),
                  THX_ERR_RANGE, "local scorer: padded volume (or its bricked copy) above 2 GiB");
    }
    int grid = 0;
    if (mode == kLocalQueue) {
        THX_TRY(c->locCounter.ensure(8 * sizeof(int)));
        grid = local_queue_grid(precision, TPAD, ns);
    }
#define THX_LOCAL_CALL(PR, NG)                                                                  \
    launch_local_prec<PR, NG>(c->stream, TPAD, nblk, mode, grid, c->locCounter.as<int>(), b.vols,  \
                          b.volsLo, b.pcls, c->idim * c->pf, c->pf, c->iCol.as<const int>(),             \
                          c->iRow.as<const int>(), c->iSig.as<const int>(),                    \
                          PR == THX_FP16 ? c->shellScale.as<const float>() : nullptr, c->K,    \
                          c->Kp, c->idim,                                                      \
                          c->Y.as<const float2>(), c->Wt.as<const float>(),                    \
                          c->Aterm.as<const double>(), b.quat, b.pRf, b.rOff, b.nR, b.trans,   \
                          b.pTf, b.tOff, b.nT, b.blocks, b.bMax, b.bWC, b.bArg, b.bWT, b.wRrel, \
                          nItemsDev)
    if (precision == THX_FP32) THX_LOCAL_CALL(THX_FP32, 4);
    else if (precision == THX_BF16 && ns == 8) THX_LOCAL_CALL(THX_BF16, 8);
    else if (precision == THX_BF16) THX_LOCAL_CALL(THX_BF16, 4);
    else if (ns == 8) THX_LOCAL_CALL(THX_FP16, 8);
    else THX_LOCAL_CALL(THX_FP16, 4);
#undef THX_LOCAL_CALL
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// FP64 merge of a particle's local blocks (one block per particle).
__global__ __launch_bounds__(256) void k_merge_local(
    const float* __restrict__ bMax, const float* __restrict__ bWC, const int* __restrict__ bArg,
    const float* __restrict__ bWT, const float* __restrict__ wRrel, const int* __restrict__ bStart,
    const int* __restrict__ rOff, const int* __restrict__ nR, const int* __restrict__ tOff,
    const int* __restrict__ nT, int affR, int affT, int affB, int wtStride, int rb,
    double* __restrict__ wC, double* __restrict__ wR, double* __restrict__ wT,
    double* __restrict__ base, int* __restrict__ bestR, int* __restrict__ bestT) {
    const int p = blockIdx.x;
    // a particle's blocks are contiguous from bStart[p] (particles themselves
    // may be listed in any order); null tables mean equal counts for all
    const int Rp = nR ? nR[p] : affR, Tp = nT ? nT[p] : affT;
    const size_t ro = rOff ? (size_t)rOff[p] : (size_t)p * affR;
    const size_t to = tOff ? (size_t)tOff[p] : (size_t)p * affT;
    const int b0 = bStart ? bStart[p] : p * affB, nb = (Rp + rb - 1) / rb;
    __shared__ double f[64];
    __shared__ double red[256];
    __shared__ int redi[256];
    double mx = -INFINITY;
    int ai = -1;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if ((double)bMax[b0 + i] > mx) { mx = bMax[b0 + i]; ai = i; }
    red[threadIdx.x] = mx;
    redi[threadIdx.x] = ai;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            const int oi = redi[threadIdx.x + s];
            if (o > red[threadIdx.x] || (o == red[threadIdx.x] && oi >= 0 &&
                                         (redi[threadIdx.x] < 0 || oi < redi[threadIdx.x]))) {
                red[threadIdx.x] = o;
                redi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double M = red[0];
    const int bb = redi[0];
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) f[i] = exp((double)bMax[b0 + i] - M);
    __syncthreads();
    if (threadIdx.x == 0) {
        double wc = 0.0;
        for (int i = 0; i < nb; i++) wc += f[i] * bWC[b0 + i];
        wC[p] = wc;
        base[p] = M;
        const int a = bb >= 0 ? bArg[b0 + bb] : -1;
        bestR[p] = a < 0 ? -1 : a / kNBL;
        bestT[p] = a < 0 ? -1 : a % kNBL;
    }
    for (int r = threadIdx.x; r < Rp; r += blockDim.x)
        wR[ro + r] = f[r / rb] * (double)wRrel[ro + r];
    for (int t = threadIdx.x; t < Tp; t += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < nb; i++) s += f[i] * bWT[(size_t)(b0 + i) * wtStride + t];
        wT[to + t] = s;
    }
}

// importance-weight ESS per particle over its selected rotations:
// w_r = wR_r / q_r (translation-marginal likelihood over the proposal density)
// Per particle over its scored selection: the effective sample size of the
// importance weights w_r = wR_r / q_r, (sum w)^2 / sum w^2, the tail depth of
// the best rotation, max_r log q_r - log q_best, and the best selected
// rotation / translation mapped to grid indices (in place).
__global__ void k_sel_stats(const double* __restrict__ wR, const float* __restrict__ rlogq,
                            const int* __restrict__ ridx, const int* __restrict__ tidx,
                            int* __restrict__ bestR, int* __restrict__ bestT, int nSelR,
                            int nSelT, int P, double* __restrict__ ess, float* __restrict__ tail) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float* lq = rlogq + (size_t)p * nSelR;
    double mx = -INFINITY;
    float lqmax = -INFINITY;
    for (int r = 0; r < nSelR; r++) {
        mx = fmax(mx, -(double)lq[r]);
        lqmax = fmaxf(lqmax, lq[r]);
    }
    double s1 = 0.0, s2 = 0.0;
    for (int r = 0; r < nSelR; r++) {
        const double w = wR[(size_t)p * nSelR + r] * exp(-(double)lq[r] - mx);
        s1 += w;
        s2 += w * w;
    }
    ess[p] = s2 > 0.0 ? s1 * s1 / s2 : 0.0;
    const int b = bestR[p], bt = bestT[p];
    tail[p] = b >= 0 ? lqmax - lq[b] : INFINITY;
    bestR[p] = b >= 0 ? ridx[(size_t)p * nSelR + b] : -1;
    bestT[p] = bt >= 0 ? tidx[(size_t)p * nSelT + bt] : -1;
}

int launch_sel_stats(hipStream_t st, const double* wR, const float* rlogq, const int* ridx,
                     const int* tidx, int* bestR, int* bestT, int nSelR, int nSelT, int P,
                     double* ess, float* tail) {
    hipLaunchKernelGGL(k_sel_stats, dim3((P + 127) / 128), dim3(128), 0, st, wR, rlogq, ridx,
                       tidx, bestR, bestT, nSelR, nSelT, P, ess, tail);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// Rows idx[i] of the particle inputs (data, CTF, sigma) copied to row i.
__global__ __launch_bounds__(256) void k_gather_particles(
    const float2* __restrict__ dat, const float* __restrict__ ctf, const float* __restrict__ sig,
    const int* __restrict__ idx, int K, float2* __restrict__ dat2, float* __restrict__ ctf2,
    float* __restrict__ sig2) {
    // blockIdx.y splits the row so a few hundred rows still fill the machine
    const size_t src = (size_t)idx[blockIdx.x] * K, dst = (size_t)blockIdx.x * K;
    for (int k = blockIdx.y * blockDim.x + threadIdx.x; k < K; k += gridDim.y * blockDim.x) {
        dat2[dst + k] = dat[src + k];
        ctf2[dst + k] = ctf[src + k];
        sig2[dst + k] = sig[src + k];
    }
}

int launch_gather_particles(thx_ctx* c, const int* idx, int n, float2* dat2, float* ctf2,
                            float* sig2) {
    if (n <= 0) return THX_OK;
    hipLaunchKernelGGL(k_gather_particles, dim3(n, 8), dim3(256), 0, c->stream,
                       c->dat.as<const float2>(), c->ctf.as<const float>(),
                       c->sig.as<const float>(), idx, c->K, dat2, ctf2, sig2);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

int launch_merge_local(thx_ctx* c, int P, int wtStride, const LocalBatch& b, double* wC,
                       double* wR, double* wT, double* base, int* bestR, int* bestT, int rb) {
    hipLaunchKernelGGL(k_merge_local, dim3(P), dim3(256), 0, c->stream, b.bMax, b.bWC, b.bArg,
                       b.bWT, b.wRrel, b.bStart, b.rOff, b.nR, b.tOff, b.nT, b.affR, b.affT,
                       b.affB, wtStride, rb, wC, wR, wT, base, bestR, bestT);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Selected-pose scorer of the grid-restricted proposal.  Particle p scores
// its nSelR selected grid rotations (rows of the shared slice table rotLo,
// BF16) against its nSelT selected grid translations (rows of the phase
// table, read from its FP16 copy S16: a unit phase in FP16 is within 2^-12
// per component, well inside the BF16 rounding of Z = Y conj(S) that
// follows); the norm term is accumulated in FP32 from the 16-bit slice values.  A
// block owns one particle x up to RB of its selected rotations x all its
// selected translations; epilogue and merge are those of the local scorer.
// With an order table, hardware block b takes the particle order[j / nbp]
// with j = xcd_logical(b): each XCD walks a contiguous run of the ordered
// particles, so the blocks resident on one XCD at a time score particles whose
// selected slice rows overlap and share its L2.  Outputs are indexed by the
// particle's own block (p nbp + rotation block) either way.
// ---------------------------------------------------------------------------
template <int TPAD, int RB>
__global__ __launch_bounds__(2 * RB) void k_score_sel(
    const uint32_t* __restrict__ rotLo, const float* __restrict__ normP,
    const uint32_t* __restrict__ S16, int Kp, const float2* __restrict__ Y,
    const float* __restrict__ Wt, const double* __restrict__ Aterm, const int* __restrict__ ridx,
    const int* __restrict__ tidx, int nSelR, int nSelT, float pRu, float pTu,
    const int* __restrict__ pmap, const int* __restrict__ order, float* __restrict__ bMax,
    float* __restrict__ bWC, int* __restrict__ bArg, float* __restrict__ bWT,
    float* __restrict__ wRrel) {
    constexpr int LD = 72;
    constexpr int MT = TPAD / 32;
    constexpr int NTH = 2 * RB, NW = NTH / 64;   // 128 rotations: 4 waves; 256: 8
    __shared__ __attribute__((aligned(16))) unsigned short As[TPAD * LD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[RB * LD];
    __shared__ int rrow[RB];
    __shared__ int trow[TPAD];
    __shared__ float Cs[RB];
    __shared__ float redMax[NW];
    __shared__ int redArg[NW];
    __shared__ float rowPart[NW][TPAD];
    __shared__ float rowW[TPAD];

    const int nbp = (nSelR + RB - 1) / RB;
    int bl = blockIdx.x;   // the particle's own block index
    if (order) {
        const int j = xcd_logical(blockIdx.x, gridDim.x);
        bl = order[j / nbp] * nbp + j % nbp;
    }
    const int p = bl / nbp, r0 = (bl % nbp) * RB;
    const int pd = pmap ? pmap[p] : p;   // particle row of Y / Wt / Aterm
    const int nr = min(RB, nSelR - r0);
    const int Tp = nSelT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hl = lane >> 5;
    // null selections mean the whole grid (rotation / translation index = position)
    if (tid < RB) rrow[tid] = tid < nr ? (ridx ? ridx[(size_t)p * nSelR + r0 + tid] : r0 + tid) : -1;
    if (tid < TPAD) trow[tid] = tid < Tp ? (tidx ? tidx[(size_t)p * nSelT + tid] : tid) : -1;
    __syncthreads();

    // Loads are 16 bytes per lane, 4 pixels: lanes l8 = tid & 7 of eight
    // consecutive lanes cover a row's 32-pixel chunk, slice rows rq + NTH/8 q
    // and phase rows rq + NTH/8 j (the particle's Y and weights at the same
    // pixels).  The norm term is taken from the 16-bit slice values, so no
    // |P|^2 table is read.  Rows past the selection read row 0 of their table:
    // their columns (rotations) and rows (translations) of the tile are masked
    // in the epilogue, so the loads need no predicate; phase rows past TPAD
    // (TPAD < NTH/8) are loaded but not staged.
    constexpr int SR = NTH / 8;                  // phase rows per pass
    constexpr int NJ = (TPAD + SR - 1) / SR;
    const int l8 = tid & 7, rq = tid >> 3;
    // per-lane 32-bit byte offsets from uniform bases (the tables stay below
    // 4 GiB): each load is a scalar base plus a vector offset, with no 64-bit
    // address arithmetic per load
    const char* bBase = reinterpret_cast<const char*>(rotLo);
    const char* sBase = reinterpret_cast<const char*>(S16);
    const char* yBase = reinterpret_cast<const char*>(Y + (size_t)pd * Kp);
    const char* wBase = reinterpret_cast<const char*>(Wt + (size_t)pd * Kp);
    uint32_t boff[4], soff[NJ];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int row = rrow[rq + (NTH / 8) * q];
        boff[q] = ((uint32_t)(row >= 0 ? row : 0) * (uint32_t)Kp + 4u * l8) * 4u;
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int tr = rq + SR * j;
        const int t = tr < TPAD ? trow[tr] : -1;
        soff[j] = ((uint32_t)(t >= 0 ? t : 0) * (uint32_t)Kp + 4u * l8) * 4u;
    }
    const uint32_t yoff = 32u * l8, woff = 16u * l8;

    floatx16 acc[MT];
#pragma unroll
    for (int a = 0; a < MT; a++) acc[a] = floatx16{0};
    float cpart[4] = {0.f, 0.f, 0.f, 0.f};
    u32x4 bv[4];
    u32x4 sv[NJ];
    float4 yv0, yv1, wv;
    auto gload = [&](int k0) {
        const uint32_t kb = (uint32_t)k0 * 4u, ks = (uint32_t)k0 * 8u;   // chunk byte offsets
#pragma unroll
        for (int q = 0; q < 4; q++) bv[q] = *reinterpret_cast<const u32x4*>(bBase + (boff[q] + kb));
#pragma unroll
        for (int j = 0; j < NJ; j++) sv[j] = *reinterpret_cast<const u32x4*>(sBase + (soff[j] + kb));
        yv0 = *reinterpret_cast<const float4*>(yBase + (yoff + ks));
        yv1 = *reinterpret_cast<const float4*>(yBase + (yoff + ks + 16u));
        wv = *reinterpret_cast<const float4*>(wBase + (woff + kb));
    };
    gload(0);
    for (int k0 = 0; k0 < Kp; k0 += kChunk) {
        // ---- norm term and B staging ----
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float w4[4] = {wv.x, wv.y, wv.z, wv.w};
            float c = 0.f;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                // |P|^2 of one packed BF16 pixel: re^2 + im^2 (exact products;
                // the element goes through a scalar: a bit_cast of a vector
                // element expression reads element 0 with this compiler)
                const uint32_t u = bv[q][e];
                const bf16x2 v = __builtin_bit_cast(bf16x2, u);
                c += w4[e] * __builtin_amdgcn_fdot2_f32_bf16(v, v, 0.f, false);
            }
            cpart[q] += c;
            *reinterpret_cast<u32x4*>(Bs + (rq + (NTH / 8) * q) * LD + 8 * l8) = bv[q];
        }
        // ---- A: Z = Y conj(S_t) ----
        const float yr[4] = {yv0.x, yv0.z, yv1.x, yv1.z}, yi[4] = {yv0.y, yv0.w, yv1.y, yv1.w};
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int tr = rq + SR * j;
            u32x4 z;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t u = sv[j][e];
                const float sr = (float)__builtin_bit_cast(_Float16, (unsigned short)(u & 0xffffu));
                const float si = (float)__builtin_bit_cast(_Float16, (unsigned short)(u >> 16));
                z[e] = pack16<THX_BF16>(yr[e] * sr + yi[e] * si, yi[e] * sr - yr[e] * si);
            }
            if (SR * NJ == TPAD || tr < TPAD)
                *reinterpret_cast<u32x4*>(As + tr * LD + 8 * l8) = z;
        }
        __syncthreads();
        // the next chunk's loads, unconditional (the last iteration reloads its
        // own chunk, unused): a conditional load would make the compiler merge
        // old and new register values with copies on every iteration
        gload(k0 + kChunk < Kp ? k0 + kChunk : k0);
        // keep the loads ahead of the MFMA steps: left free, the scheduler
        // interleaves the norm term's dot products with the MFMAs and sinks
        // the loads (which overwrite those registers) to the end of the
        // iteration, where nothing hides their latency
        __builtin_amdgcn_sched_barrier(0);
        const int nrow = 32 * wave + l32;
#pragma unroll
        for (int kk = 0; kk < 64; kk += 16) {
            const int kc = kk + 8 * hl;
            const uint4 rb4 = *reinterpret_cast<const uint4*>(Bs + nrow * LD + kc);
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const uint4 ra = *reinterpret_cast<const uint4*>(As + (32 * mt + l32) * LD + kc);
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8, ra), __builtin_bit_cast(bf16x8, rb4), acc[mt], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // ---- norm term per rotation: sum over the eight lanes of a row ----
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float v = cpart[q];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        if (l8 == 0) Cs[rq + (NTH / 8) * q] = v;
    }
    __syncthreads();

    const float a = (float)Aterm[pd];
    const int ncol = 32 * wave + l32;
    const bool cok = ncol < nr;
    const float cterm = Cs[ncol];
    float lmax = -INFINITY;
    int larg = -1;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            float v = -INFINITY;
            if (cok && t < Tp) v = a + cterm - 2.f * acc[mt][i];
            acc[mt][i] = v;
            if (v > lmax) { lmax = v; larg = ncol * TPAD + t; }
        }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(lmax, off);
        const int oa = __shfl_xor(larg, off);
        if (o > lmax || (o == lmax && oa >= 0 && (larg < 0 || oa < larg))) { lmax = o; larg = oa; }
    }
    if (lane == 0) { redMax[wave] = lmax; redArg[wave] = larg; }
    __syncthreads();
    float M = redMax[0];
    int arg = redArg[0];
    for (int w = 1; w < NW; w++)
        if (redMax[w] > M || (redMax[w] == M && redArg[w] >= 0 && (arg < 0 || redArg[w] < arg))) {
            M = redMax[w];
            arg = redArg[w];
        }
    const bool live = M > -INFINITY;
    constexpr int NV = MT * 16 < 32 ? 32 : MT * 16;
    float rowv[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) rowv[j] = 0.f;
    const float pr = cok ? pRu : 0.f;
    float cs = 0.f;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            const float w = live ? __expf(acc[mt][i] - M) : 0.f;
            cs += w * (t < Tp ? pTu : 0.f);
            rowv[mt * 16 + i] = w * pr;
        }
    cs += __shfl_xor(cs, 32);
    if (hl == 0 && cok) wRrel[(size_t)p * nSelR + r0 + ncol] = cs;
    constexpr int NG = NV / 32;
#pragma unroll
    for (int g = 0; g < NG; g++) {
#pragma unroll
        for (int s = 16; s >= 1; s >>= 1) {
            const bool up = (l32 & s) != 0;
#pragma unroll
            for (int j = 0; j < s; j++) {
                const float send = up ? rowv[g * 32 + j] : rowv[g * 32 + j + s];
                const float keep = up ? rowv[g * 32 + j + s] : rowv[g * 32 + j];
                rowv[g * 32 + j] = keep + __shfl_xor(send, s);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int jv = g * 32 + l32;
        const int mt = jv >> 4, i = jv & 15;
        if (mt < MT) {
            const int t = 32 * mt + 8 * (i >> 2) + 4 * hl + (i & 3);
            rowPart[wave][t] = rowv[g * 32];
        }
    }
    __syncthreads();
    if (tid < TPAD) {
        float rsum = 0.f;
        for (int w = 0; w < NW; w++) rsum += rowPart[w][tid];
        bWT[(size_t)bl * TPAD + tid] = rsum;
        rowW[tid] = tid < Tp ? rsum * pTu : 0.f;
    }
    __syncthreads();
    if (tid == 0) {
        float wc = 0.f;
        for (int t = 0; t < TPAD; t++) wc += rowW[t];
        bMax[bl] = M;
        bWC[bl] = wc;
        bArg[bl] = arg < 0 ? -1 : ((r0 + arg / TPAD) * kNBL + arg % TPAD);
    }
}

int launch_score_sel(thx_ctx* c, int TPAD, int RB, int nblk, const int* ridx, const int* tidx,
                     int nSelR, int nSelT, float pRu, float pTu, const int* pmap,
                     const int* order, const LocalBatch& b) {
#define THX_SEL_ARGS c->rotLo.as<const uint32_t>(), c->normP.as<const float>(),            \
                     c->S16.as<const uint32_t>(), c->Kp, c->Y.as<const float2>(),          \
                     c->Wt.as<const float>(), c->Aterm.as<const double>(), ridx, tidx,     \
                     nSelR, nSelT, pRu, pTu, pmap, order, b.bMax, b.bWC, b.bArg, b.bWT,   \
                     b.wRrel
    THX_CHECK(RB == 128 || RB == 256, THX_ERR_ARG, "score_sel: rotation block must be 128 or 256");
    // the loads address the slice and phase tables with 32-bit byte offsets
    THX_CHECK(c->rotLo.bytes <= (1ull << 32) && c->S16.bytes <= (1ull << 32), THX_ERR_RANGE,
              "score_sel: slice or phase table above 4 GiB");
    THX_CHECK(c->S16.bytes >= sizeof(uint32_t) * (size_t)c->TpadS * c->Kp, THX_ERR_STATE,
              "score_sel: FP16 phase table not built");
    if (RB == 128) {
        if (TPAD == 32)
            hipLaunchKernelGGL((k_score_sel<32, 128>), dim3(nblk), dim3(256), 0, c->stream, THX_SEL_ARGS);
        else if (TPAD == 64)
            hipLaunchKernelGGL((k_score_sel<64, 128>), dim3(nblk), dim3(256), 0, c->stream, THX_SEL_ARGS);
        else
            hipLaunchKernelGGL((k_score_sel<128, 128>), dim3(nblk), dim3(256), 0, c->stream, THX_SEL_ARGS);
    } else {
        if (TPAD == 32)
            hipLaunchKernelGGL((k_score_sel<32, 256>), dim3(nblk), dim3(512), 0, c->stream, THX_SEL_ARGS);
        else if (TPAD == 64)
            hipLaunchKernelGGL((k_score_sel<64, 256>), dim3(nblk), dim3(512), 0, c->stream, THX_SEL_ARGS);
        else
            hipLaunchKernelGGL((k_score_sel<128, 256>), dim3(nblk), dim3(512), 0, c->stream, THX_SEL_ARGS);
    }
#undef THX_SEL_ARGS
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Direct-form scores of one particle (checking aid): one block per (r, t),
// dvp = sum_k sig |X - ctf S P|^2 in FP32 (the kernel_logDataVS form).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_score_one(
    const float2* __restrict__ dat, const float* __restrict__ ctf,
    const float* __restrict__ sig, const float2* __restrict__ rotP,
    const float2* __restrict__ S, int p, int K, int Kp, int T, float* __restrict__ out) {
    const int r = blockIdx.x / T, t = blockIdx.x % T;
    float acc = 0.f;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float2 pr = rotP[(size_t)r * Kp + k], s = S[(size_t)t * Kp + k];
        const float qr = s.x * pr.x - s.y * pr.y, qi = s.x * pr.y + s.y * pr.x;
        const size_t i = (size_t)p * K + k;
        const float c = ctf[i];
        const float er = dat[i].x - c * qr, ei = dat[i].y - c * qi;
        acc += (er * er + ei * ei) * sig[i];
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

int launch_score_one(thx_ctx* c, int cls, int p, float* dOut) {
    (void)cls;
    hipLaunchKernelGGL(k_score_one, dim3(c->R * c->T), dim3(256), 0, c->stream,
                       c->dat.as<const float2>(), c->ctf.as<const float>(),
                       c->sig.as<const float>(), c->rotP.as<const float2>(),
                       c->S.as<const float2>(), p, c->K, c->Kp, c->T, dOut);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Synthetic workloads.
// ---------------------------------------------------------------------------

// Analytic FT of the blob sum on the padded half-complex grid:
// F(m) = (2 pi)^{3/2} s^3 exp(-2 pi^2 s^2 |m|^2 / V^2) sum_b exp(-2 pi i m.c_b / V).
__global__ __launch_bounds__(256) void k_synth_volume(
    const double* __restrict__ centres, const double* __restrict__ amps, int nBlob, double sigma,
    int V, float2* __restrict__ vol) {
    const int nColFT = V / 2 + 1;
    const size_t n = (size_t)nColFT * V * V;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double cs[3 * 256];
    __shared__ double am[256];   // blob amplitudes (1 without amps)
    double re = 0.0, im = 0.0;
    int mx = 0, my = 0, mz = 0;
    if (idx < n) {
        mx = (int)(idx % nColFT);
        my = (int)((idx / nColFT) % V);
        mz = (int)(idx / ((size_t)nColFT * V));
        if (my >= V / 2) my -= V;
        if (mz >= V / 2) mz -= V;
    }
    for (int b0 = 0; b0 < nBlob; b0 += 256) {
        __syncthreads();
        for (int i = threadIdx.x; i < 3 * 256; i += blockDim.x)
            cs[i] = (b0 * 3 + i < 3 * nBlob) ? centres[b0 * 3 + i] : 0.0;
        for (int i = threadIdx.x; i < 256; i += blockDim.x)
            am[i] = (b0 + i < nBlob) ? (amps ? amps[b0 + i] : 1.0) : 0.0;
        __syncthreads();
        const int nb = min(256, nBlob - b0);
        for (int b = 0; b < nb; b++) {
            const double ph = -2.0 * (mx * cs[3 * b] + my * cs[3 * b + 1] + mz * cs[3 * b + 2]) / V;
            double sn, c;
            sincospi(ph, &sn, &c);
            re += am[b] * c;
            im += am[b] * sn;
        }
    }
    if (idx < n) {
        const double m2 = (double)mx * mx + (double)my * my + (double)mz * mz;
        const double g = pow(2 * M_PI, 1.5) * sigma * sigma * sigma *
                         exp(-2 * M_PI * M_PI * sigma * sigma * m2 / ((double)V * V));
        vol[idx] = make_float2((float)(g * re), (float)(g * im));
    }
}

int launch_synth_volume(thx_ctx* c, int cls, int idim, int pf, const double* dCentres,
                        int nBlob, double sigma, const double* dAmps) {
    const int V = idim * pf;
    const size_t n = (size_t)(V / 2 + 1) * V * V;
    hipLaunchKernelGGL(k_synth_volume, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       dCentres, dAmps, nBlob, sigma, V, c->vol[cls].as<float2>());
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// Particle signal CTF (.) S_t (.) P_r at the pixel set.  params per particle:
// q(4), shift(2), dU, dV, theta (9 doubles).  energy[p] += |signal|^2.
__global__ __launch_bounds__(256) void k_synth_signal(
    const float2* __restrict__ vol, int V, int pf, const double* __restrict__ params,
    const int* __restrict__ iCol, const int* __restrict__ iRow, int K, int N,
    float voltage, float Cs, float ampC, float pixelSize, float2* __restrict__ dat,
    float* __restrict__ ctfOut, double* __restrict__ energy) {
    const int p = blockIdx.x;   // particles on x, the grid dimension with the largest limit
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (k < K) {
        const double* q = params + 9 * (size_t)p;
        double m[3][2];
        quat_cols01(q, m);
        const double c0 = (double)(iCol[k] * pf), c1 = (double)(iRow[k] * pf);
        const float x = (float)(m[0][0] * c0 + m[0][1] * c1);
        const float y = (float)(m[1][0] * c0 + m[1][1] * c1);
        const float z = (float)(m[2][0] * c0 + m[2][1] * c1);
        const float2 v = sample_trilinear(vol, V, x, y, z);
        const float rCol = (float)q[4] / N, rRow = (float)q[5] / N;
        const float phase = (float)(2 * M_PI) * (iCol[k] * rCol + iRow[k] * rRow);
        const float sr = cosf(-phase), si = sinf(-phase);
        const float lambda = (float)(12.2643247 / sqrt((double)voltage * (1 + voltage * 0.978466e-6)));
        const float cv = ctf_value(iCol[k], iRow[k], pixelSize, lambda, (float)q[6], (float)q[7],
                                   (float)q[8], Cs, ampC, N);
        const float qr = sr * v.x - si * v.y, qi = sr * v.y + si * v.x;
        const float2 s = make_float2(cv * qr, cv * qi);
        dat[(size_t)p * K + k] = s;
        ctfOut[(size_t)p * K + k] = cv;
        e = (double)s.x * s.x + (double)s.y * s.y;
    }
    __shared__ double red[256];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    // one partial per block, summed in order on the host (deterministic)
    if (threadIdx.x == 0) energy[(size_t)p * gridDim.y + blockIdx.y] = red[0];
}

int launch_synth_signal(thx_ctx* c, int cls, int P, const double* dParams, double* dEnergy,
                        double voltage, double Cs, double ampContrast, double pixelSize) {
    dim3 grid(P, (c->K + 255) / 256);
    hipLaunchKernelGGL(k_synth_signal, grid, dim3(256), 0, c->stream, c->vol[cls].as<const float2>(),
                       c->vdim[cls], c->pf, dParams, c->iCol.as<const int>(),
                       c->iRow.as<const int>(), c->K, c->idim, (float)voltage, (float)Cs,
                       (float)ampContrast, (float)pixelSize, c->dat.as<float2>(),
                       c->ctf.as<float>(), dEnergy);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// Complex white noise of per-component standard deviation noiseSigma
// (Box-Muller on Philox words), and sigRcp = -1 / (2 noiseSigma^2).
__global__ __launch_bounds__(256) void k_synth_noise(float2* __restrict__ dat,
                                                     float* __restrict__ sig, int P, int K,
                                                     double noiseSigma, uint64_t seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * K) return;
    uint32_t w[4];
    Philox::draw(seed, (uint32_t)i, (uint32_t)(i >> 32), kStreamNoise, w);
    const double u1 = Philox::u01(w[0], w[1]), u2 = Philox::u01(w[2], w[3]);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    float2 d = dat[i];
    d.x += (float)(noiseSigma * rad * c);
    d.y += (float)(noiseSigma * rad * s);
    dat[i] = d;
    sig[i] = (float)(-1.0 / (2.0 * noiseSigma * noiseSigma));
}

int launch_synth_noise(thx_ctx* c, int P, double noiseSigma, uint64_t seed) {
    const size_t n = (size_t)P * c->K;
    hipLaunchKernelGGL(k_synth_noise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->dat.as<float2>(), c->sig.as<float>(), P, c->K, noiseSigma, seed);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

}  // namespace thx
