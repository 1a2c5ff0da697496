// Systematic resampling on device buffers without any allocation, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The computation is that of thx_pf_resample (k_resample in
// thx_pf.hip); the only device memory beyond the caller's inputs and outputs
// is the per-particle offsets and checked counts, which live in the
// file-scope array `scratch` below.  Each GPU holds its own copy of that array
// (2 MiB)
// from the first call on it until the process ends, and every call on that GPU
// shares it: calls that overlap on one GPU race on the offsets.

#include <hip/hip_runtime.h>

#include <mutex>

#include "thx_internal.hpp"

namespace thx {

// Largest batch of the no-allocation entry point.  With counts clamped to
// [0, kResampleMaxN] the offsets stay below 2^30, inside int range.
constexpr int kNoAllocMaxP = THX_NOALLOC_MAX_P;
constexpr size_t kScratchInts = 4 * (size_t)(kNoAllocMaxP + 1);

// offsets of the input samples [P + 1], of the output samples [P + 1], and
// the checked counts [P + 1] of each (0 for a particle whose counts are not
// both in [1, kResampleMaxN], which k_resample then skips)
static __device__ __align__(256) int scratch[kScratchInts];

// Device address of `scratch` on each device, taken once per device with that
// device current (hipGetSymbolAddress).  Host code uses these addresses,
// never the array's name, which is only a host-side stand-in.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int* g_scratch_ptr[kMaxDevices] = {};

static int scratch_on(int device, int** out) {
    THX_CHECK(device >= 0 && device < kMaxDevices, THX_ERR_ARG, "pf_resample_noalloc: bad device");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        THX_HIP(hipSetDevice(device));
        THX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        THX_CHECK(ptr != nullptr, THX_ERR_HIP, "pf_resample_noalloc: no scratch address");
        g_scratch_ptr[device] = static_cast<int*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return THX_OK;
}

// Exclusive prefix sums of the input and output counts (each clamped to
// [0, kResampleMaxN]) into offIn[0..P] and offOut[0..P], and the checked
// counts nInV / nOutV (the counts, or 0 when either is outside
// [1, kResampleMaxN]), in one block of 1024 threads: each thread sums a
// contiguous range, the range sums are scanned in LDS, then each range is
// written.  Every entry used by the call is rewritten, so nothing from the
// previous call survives.
__global__ __launch_bounds__(1024) void k_count_offsets(const int* __restrict__ nIn,
                                                        const int* __restrict__ nOut, int P,
                                                        int* __restrict__ offIn,
                                                        int* __restrict__ offOut,
                                                        int* __restrict__ nInV,
                                                        int* __restrict__ nOutV) {
    __shared__ int sIn[1024], sOut[1024];
    const int t = threadIdx.x, per = (P + 1023) / 1024;
    const int b0 = min(P, t * per), b1 = min(P, b0 + per);
    int a = 0, b = 0;
    for (int p = b0; p < b1; p++) {
        a += min(max(nIn[p], 0), kResampleMaxN);
        b += min(max(nOut[p], 0), kResampleMaxN);
    }
    sIn[t] = a;
    sOut[t] = b;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 1024 range sums
    for (int d = 1; d < 1024; d <<= 1) {
        const int x = t >= d ? sIn[t - d] : 0, y = t >= d ? sOut[t - d] : 0;
        __syncthreads();
        sIn[t] += x;
        sOut[t] += y;
        __syncthreads();
    }
    int oi = t > 0 ? sIn[t - 1] : 0, oo = t > 0 ? sOut[t - 1] : 0;
    for (int p = b0; p < b1; p++) {
        const int ni = nIn[p], no = nOut[p];
        const bool ok = ni >= 1 && ni <= kResampleMaxN && no >= 1 && no <= kResampleMaxN;
        offIn[p] = oi;
        offOut[p] = oo;
        nInV[p] = ok ? ni : 0;
        nOutV[p] = ok ? no : 0;
        oi += min(max(ni, 0), kResampleMaxN);
        oo += min(max(no, 0), kResampleMaxN);
    }
    if (t == 1023) {
        offIn[P] = sIn[1023];
        offOut[P] = sOut[1023];
    }
}

}  // namespace thx

using namespace thx;

extern "C" int thx_pf_resample_noalloc(thx_ctx* c, const double* w, const double* u,
                                       const int* nIn, const int* nOut, int P, uint64_t seed,
                                       const double* u0, int* idx, double* wNew, double* ess) {
    clear_error();
    THX_CHECK(c && w && u && nIn && nOut && idx && wNew && P > 0, THX_ERR_ARG,
              "pf_resample_noalloc: bad argument");
    THX_CHECK(P <= kNoAllocMaxP, THX_ERR_RANGE,
              "pf_resample_noalloc: more particles than the scratch array holds "
              "(THX_NOALLOC_MAX_P)");
    int* base = nullptr;
    THX_TRY(scratch_on(c->device, &base));
    THX_HIP(hipSetDevice(c->device));
    int* offIn = base;
    int* offOut = base + (P + 1);
    int* nInV = base + 2 * (P + 1);
    int* nOutV = base + 3 * (P + 1);
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_count_offsets, dim3(1), dim3(1024), 0, st, nIn, nOut, P, offIn, offOut,
                       nInV, nOutV);
    THX_HIP(hipGetLastError());
    return pf_resample(st, w, u, offIn, nInV, offOut, nOutV, P, seed, 0, u0, idx, wNew, ess);
}
