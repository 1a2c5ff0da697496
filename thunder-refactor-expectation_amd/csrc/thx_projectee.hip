// Projectee preparation on the GPU: Projector::setProjectee(Volume)
// (src/Projector.cpp:123-148) with the LINEAR-interpolation grid correction
// (Projector::gridCorrection, src/Projector.cpp:524-606, active under
// PROJECTOR_CORRECT_CONVOLUTION_KERNEL, include/Config.h:125).
//
//   1. inverse real FFT of the N^3 half-complex map, scaled by 1 / N^3
//      (FFT::bw(Volume&), src/FFT.cpp:201-230);
//   2. zero padding to V^3 = (pf N)^3 around the origin, negative indices
//      wrapped at both sizes (VOL_PAD_RL, include/Image/ImageFunctions.h:176-192;
//      Volume::iRL, include/Image/Volume.h:519-527);
//   3. division by TIK_RL(r / (pf V)) = j0(pi r / (pf V))^2, r = |(i, j, k)| of
//      the padded voxel and V the padded size (src/Functions/Functions.cpp:217-220);
//   4. forward real FFT of the padded map, unnormalised, into the
//      (V/2 + 1) V V half-complex projectee (FFT::fw(Volume&), src/FFT.cpp:172-199).
//
// The two FFTs are plain library transforms (hipFFT, FP32 like the reference's
// single-precision FFTW build); steps 2 and 3 are one hand-written kernel that
// reads the N^3 map once and writes every padded voxel once.
#include <hipfft/hipfft.h>

#include <cmath>

#include "thx_internal.hpp"

namespace thx {

// dst[V^3] (x fastest): padded, scaled, grid-corrected copy of src[N^3].
// Block (x quarter-row, y, z): each lane writes 4 consecutive x voxels with
// one 16-byte store; only the N^3 voxels of the map read anything.
__global__ __launch_bounds__(64) void k_pad_correct(const float* __restrict__ src, int N, int V,
                                                     int pf, float* __restrict__ dst) {
    const int x0 = 4 * (blockIdx.x * 64 + threadIdx.x);
    if (x0 >= V) return;
    const int y = blockIdx.y, z = blockIdx.z;
    // signed coordinates of the padded voxel
    const int j = y < V / 2 ? y : y - V, k = z < V / 2 ? z : z - V;
    const int h = N / 2;
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (j >= -h && j < h && k >= -h && k < h) {
        const int sj = j >= 0 ? j : j + N, sk = k >= 0 ? k : k + N;
        const float* srow = src + ((size_t)sk * N + sj) * N;
        const double scale = 1.0 / ((double)N * N * N), w = M_PI / ((double)pf * V);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int x = x0 + e;
            const int i = x < V / 2 ? x : x - V;
            if (i >= -h && i < h) {
                const double a = w * sqrt((double)(i * i + j * j + k * k));
                const double j0 = a == 0.0 ? 1.0 : sin(a) / a;
                v[e] = srow[i >= 0 ? i : i + N] * (float)scale / (float)(j0 * j0);
            }
        }
    }
    *reinterpret_cast<float4*>(dst + ((size_t)z * V + y) * V + x0) = make_float4(v[0], v[1], v[2], v[3]);
}

static int fft_status(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) {
        set_error(std::string("set_projectee: ") + what + " failed (hipFFT status " +
                  std::to_string((int)r) + ")");
        return THX_ERR_HIP;
    }
    return THX_OK;
}

void projectee_release(thx_ctx* c) {
    if (c->prjPlanBw) (void)hipfftDestroy(static_cast<hipfftHandle>(c->prjPlanBw));
    if (c->prjPlanFw) (void)hipfftDestroy(static_cast<hipfftHandle>(c->prjPlanFw));
    c->prjPlanBw = c->prjPlanFw = nullptr;
    c->prjN = 0;
    c->prjIn.release();
    c->prjReal.release();
    c->prjPad.release();
}

}  // namespace thx

using namespace thx;

int thx_ctx_set_projectee(thx_ctx* c, int cls, const float* volFT, int N, int pf,
                          float* msOut) {
    clear_error();
    THX_CHECK(c && volFT && cls >= 0, THX_ERR_ARG, "set_projectee: bad argument");
    THX_CHECK(N >= 4 && N % 2 == 0 && pf >= 1 && (pf * N) % 4 == 0 && pf * N <= 1024,
              THX_ERR_ARG, "set_projectee: N must be even, >= 4, pf * N a multiple of 4 and <= 1024");
    THX_HIP(hipSetDevice(c->device));
    const int V = pf * N;
    const size_t nIn = (size_t)(N / 2 + 1) * N * N, nReal = (size_t)N * N * N;
    const size_t nPad = (size_t)V * V * V, nOut = (size_t)(V / 2 + 1) * V * V;
    hipStream_t st = c->stream;
    if (c->prjN != N || c->prjPf != pf) {
        projectee_release(c);
        hipfftHandle bw = nullptr, fw = nullptr;
        size_t work = 0;
        THX_TRY(fft_status(hipfftCreate(&bw), "plan create"));
        c->prjPlanBw = bw;
        THX_TRY(fft_status(hipfftCreate(&fw), "plan create"));
        c->prjPlanFw = fw;
        THX_TRY(fft_status(hipfftMakePlan3d(bw, N, N, N, HIPFFT_C2R, &work), "inverse plan"));
        THX_TRY(fft_status(hipfftMakePlan3d(fw, V, V, V, HIPFFT_R2C, &work), "forward plan"));
        c->prjN = N;
        c->prjPf = pf;
    }
    hipfftHandle bw = static_cast<hipfftHandle>(c->prjPlanBw);
    hipfftHandle fw = static_cast<hipfftHandle>(c->prjPlanFw);
    THX_TRY(fft_status(hipfftSetStream(bw, st), "set stream"));
    THX_TRY(fft_status(hipfftSetStream(fw, st), "set stream"));
    if ((int)c->vol.size() <= cls) {
        c->vol.resize(cls + 1);
        c->vdim.resize(cls + 1, 0);
    }
    THX_TRY(c->prjIn.ensure(sizeof(float2) * nIn));
    THX_TRY(c->prjReal.ensure(sizeof(float) * nReal));
    THX_TRY(c->prjPad.ensure(sizeof(float) * nPad));
    THX_TRY(c->vol[cls].ensure(sizeof(float2) * nOut));
    c->vdim[cls] = V;
    invalidate_vol_lo(c, cls);
    THX_HIP(hipMemcpyAsync(c->prjIn.p, volFT, sizeof(float2) * nIn, hipMemcpyHostToDevice, st));
    THX_HIP(hipEventRecord(c->ev[0], st));
    // the inverse transform may overwrite its input: prjIn is scratch
    THX_TRY(fft_status(hipfftExecC2R(bw, c->prjIn.as<hipfftComplex>(), c->prjReal.as<hipfftReal>()),
                       "inverse transform"));
    THX_HIP(hipEventRecord(c->ev[1], st));
    hipLaunchKernelGGL(k_pad_correct, dim3((unsigned)((V / 4 + 63) / 64), V, V), dim3(64), 0, st,
                       c->prjReal.as<const float>(), N, V, pf, c->prjPad.as<float>());
    THX_HIP(hipGetLastError());
    THX_HIP(hipEventRecord(c->ev[2], st));
    THX_TRY(fft_status(hipfftExecR2C(fw, c->prjPad.as<hipfftReal>(), c->vol[cls].as<hipfftComplex>()),
                       "forward transform"));
    THX_HIP(hipEventRecord(c->ev[3], st));
    THX_HIP(hipStreamSynchronize(st));
    c->nev = 0;   // the events now time this call, not the last scoring call
    c->evNames.clear();
    if (msOut) {
        // [0] inverse FFT, [1] pad + correction kernel, [2] forward FFT
        for (int s = 0; s < 3; s++) THX_HIP(hipEventElapsedTime(msOut + s, c->ev[s], c->ev[s + 1]));
    }
    return THX_OK;
}
