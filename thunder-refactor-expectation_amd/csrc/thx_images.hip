// Particle images to the packed Fourier data of the pixel set, on the GPU:
// the producer side of the expectation (Optimiser::initImg,
// src/Optimiser.cpp:4596-4794, without the file reading).
//
//   1. per image, the mean and sample standard deviation of the background,
//      the pixels with centred coordinates i^2 + j^2 > r^2, and the image
//      made (x - mean) / stddev (Optimiser::substractBgImg,
//      src/Optimiser.cpp:4917-4951; bgMeanStddev,
//      src/Image/ImageFunctions.cpp:607-620, GSL's sample standard deviation);
//      the image is written in the stored layout, its centre at index (0, 0),
//      as ImageFile's read puts it (MESH_IMAGE_INDEX,
//      include/Image/ImageFile.h:199-200, 229-244);
//   2. the unnormalised forward real FFT of every image (Optimiser::fwImg,
//      src/Optimiser.cpp:5003-5013; FFT::fwExecutePlan(Image&),
//      src/FFT.cpp:450-458), as one batched hipFFT transform (FP32, like the
//      reference's single-precision FFTW build);
//   3. the pixel set gathered image-major into the context's data array,
//      dat[p][k] = FT_p[iFTHalf(iCol[k], iRow[k])] (Optimiser::allocPreCal,
//      src/Optimiser.cpp:8005-8016; Image::iFTHalf, include/Image/Image.h:426-432).
//
// statImg (src/Optimiser.cpp:4796-4895) then measures the mean over the images
// of their background standard deviation about 0 and normaliseImg
// (:4987-5001) scales every image by its inverse: step 1 sums each written
// image's background squares, k_image_scale forms the scale, and the packing
// applies it to the transformed data (the transform is linear).  On request step 1 also
// writes the soft-masked image (maskImg with zeroMask, src/Optimiser.cpp:
// 4953-4985, and reMaskImg, :6053-6105: the image times the raised-cosine
// mask, zero beyond it), the one the expectation packs (allocPreCal(mask =
// true), src/Optimiser.cpp:637, 1457); the unmasked spectra stay resident for
// the initial sigma, which reads the unmasked images (OPTIMISER_SIGMA_MASK
// off, include/Config.h:169).
#include <hipfft/hipfft.h>

#include <algorithm>

#include "thx_internal.hpp"

namespace thx {

// Soft mask of radius r with a raised-cosine edge of kMaskEdge pixels at the
// centred pixel (i, j), as softMask(Image&, r, ew) builds it
// (src/Functions/Mask.cpp:333-350): 1 inside r, 0.5 + 0.5 cos(pi (u - r) / ew)
// on the edge, 0 beyond r + ew, u = |(i, j)| in single precision like the
// reference's RFLOAT image.
constexpr int kMaskEdge = 6;   // EDGE_WIDTH_RL, include/Macro.h:82
__device__ inline float soft_mask(int i, int j, float r) {
    const float u = (float)sqrt((double)(i * i + j * j));
    if (u > r + (float)kMaskEdge) return 0.f;
    if (u >= r) return (float)(0.5 + 0.5 * cos((double)((u - r) / (float)kMaskEdge) * M_PI));
    return 1.f;
}

// The mask of every pixel in the file layout (centre at (N/2, N/2)), built
// once per call and read by every image's block (it stays in L2).
__global__ void k_mask_table(int N, float r, float* __restrict__ table) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    table[e] = soft_mask(e % N - N / 2, e / N - N / 2, r);
}

// One block per image.  Background statistics in FP64 (two passes: the mean,
// then the sum of squared deviations), then the normalised image written
// with its centre at the origin.
__global__ __launch_bounds__(256) void k_img_normalise(const float* __restrict__ src, int N,
                                                       float r2, float* __restrict__ dst,
                                                       const float* __restrict__ mask,
                                                       float* __restrict__ dstMasked,
                                                       double2* __restrict__ bgStat) {
    const float* s = src + (size_t)blockIdx.x * N * N;
    float* d = dst + (size_t)blockIdx.x * N * N;
    const int h = N / 2;
    __shared__ double rs[256];
    __shared__ int rc[256];
    double sum = 0.0;
    int cnt = 0;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int i = e % N - h, j = e / N - h;   // centred coordinates of file pixel e
        if ((float)(i * i + j * j) > r2) {
            sum += s[e];
            cnt++;
        }
    }
    rs[threadIdx.x] = sum;
    rc[threadIdx.x] = cnt;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            rs[threadIdx.x] += rs[threadIdx.x + st];
            rc[threadIdx.x] += rc[threadIdx.x + st];
        }
        __syncthreads();
    }
    const double mean = rs[0] / rc[0];
    const int n = rc[0];
    __syncthreads();
    double ss = 0.0;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int i = e % N - h, j = e / N - h;
        if ((float)(i * i + j * j) > r2) {
            const double v = s[e] - mean;
            ss += v * v;
        }
    }
    rs[threadIdx.x] = ss;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) rs[threadIdx.x] += rs[threadIdx.x + st];
        __syncthreads();
    }
    const double sd = sqrt(rs[0] / (n - 1));
    const float m = (float)mean, inv = (float)(1.0 / sd);
    float* dm = dstMasked ? dstMasked + (size_t)blockIdx.x * N * N : nullptr;
    double ss2 = 0.0;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int x = e % N, y = e / N;
        const size_t o = (size_t)((y + h) % N) * N + (x + h) % N;
        const float v = (s[e] - m) * inv;
        d[o] = v;
        if (dm) dm[o] = v * mask[e];
        if ((float)((x - h) * (x - h) + (y - h) * (y - h)) > r2) ss2 += (double)v * v;
    }
    __syncthreads();
    rs[threadIdx.x] = ss2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) rs[threadIdx.x] += rs[threadIdx.x + st];
        __syncthreads();
    }
    // statImg's bgStddev(0, image, r) of the written image: its background
    // sum of squares and count
    if (threadIdx.x == 0) bgStat[blockIdx.x] = make_double2(rs[0], (double)n);
}

// Block-wide FP64 sum over NT threads (NT / 64 waves): wave shuffles, then
// one value per wave through LDS, added in wave order.  Every thread returns
// the total.
template <int NT = 256>
__device__ inline double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();   // red may still be read by a previous call
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) t += red[w];
    return t;
}

// The same normalisation with the image held in registers: V4 float4s per
// thread of NT cover N * N <= 4 * NT * V4 pixels (N a multiple of 4), so each
// image is read from HBM once.  A group of four x-consecutive pixels starts at a
// multiple of 4.  When N / 2 is a multiple of 4 the moved group stays
// contiguous and leaves in one 16-byte store; when N / 2 is 2 more than a
// multiple of 4 (N = 100, 180, ...) the group that the centre shift moves to
// column N - 2 wraps within its row: its first two pixels go to columns
// N - 2 and N - 1, its last two to columns 0 and 1, as two 8-byte stores.
template <int V4, int NT>
__global__ __launch_bounds__(NT) void k_img_normalise_reg(const float* __restrict__ src, int N,
                                                           float r2, float* __restrict__ dst,
                                                           const float* __restrict__ mask,
                                                           float* __restrict__ dstMasked,
                                                           double2* __restrict__ bgStat) {
    const float4* s = reinterpret_cast<const float4*>(src + (size_t)blockIdx.x * N * N);
    float* d = dst + (size_t)blockIdx.x * N * N;
    const int h = N / 2, n4 = N * N / 4;
    __shared__ double red[NT / 64];
    float4 x[V4];
#pragma unroll
    for (int v = 0; v < V4; v++) {
        const int e4 = threadIdx.x + NT * v;
        x[v] = e4 < n4 ? s[e4] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // background membership of the 4 V4 pixels, bit 4 v + c
    unsigned long long bgBits[(4 * V4 + 63) / 64] = {};
    double sum = 0.0;
    int cnt = 0;
#pragma unroll
    for (int v = 0; v < V4; v++) {
        const int e4 = threadIdx.x + NT * v;
        if (e4 < n4) {
            const int e = 4 * e4, y = e / N, x0 = e - y * N;
            const int j = y - h;
            const float xs[4] = {x[v].x, x[v].y, x[v].z, x[v].w};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int i = x0 + c - h;
                if ((float)(i * i + j * j) > r2) {
                    sum += xs[c];
                    cnt++;
                    bgBits[(4 * v + c) >> 6] |= 1ull << ((4 * v + c) & 63);
                }
            }
        }
    }
    const double n = block_sum<NT>((double)cnt, red);
    const double mean = block_sum<NT>(sum, red) / n;
    double ss = 0.0;
#pragma unroll
    for (int v = 0; v < V4; v++) {
        const float xs[4] = {x[v].x, x[v].y, x[v].z, x[v].w};
#pragma unroll
        for (int c = 0; c < 4; c++)
            if ((bgBits[(4 * v + c) >> 6] >> ((4 * v + c) & 63)) & 1ull) {
                const double dv = xs[c] - mean;
                ss += dv * dv;
            }
    }
    const double sd = sqrt(block_sum<NT>(ss, red) / (n - 1.0));
    const float m = (float)mean, inv = (float)(1.0 / sd);
    double ss2 = 0.0;   // background sum of squares of the written image
#pragma unroll
    for (int v = 0; v < V4; v++) {
        const int e4 = threadIdx.x + NT * v;
        if (e4 < n4) {
            const int e = 4 * e4, y = e / N, x0 = e - y * N;
            const int yo = y + h < N ? y + h : y + h - N, xo = x0 + h < N ? x0 + h : x0 + h - N;
            const float4 o = make_float4((x[v].x - m) * inv, (x[v].y - m) * inv,
                                         (x[v].z - m) * inv, (x[v].w - m) * inv);
            const float os[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
            for (int c = 0; c < 4; c++)
                if ((bgBits[(4 * v + c) >> 6] >> ((4 * v + c) & 63)) & 1ull)
                    ss2 += (double)os[c] * os[c];
            float* row = d + (size_t)yo * N;
            float* rowM = dstMasked ? dstMasked + (size_t)blockIdx.x * N * N + (size_t)yo * N
                                    : nullptr;
            float4 om = o;
            if (dstMasked) {
                const float4 w = reinterpret_cast<const float4*>(mask)[e4];
                om = make_float4(o.x * w.x, o.y * w.y, o.z * w.z, o.w * w.w);
            }
            if (xo + 4 <= N) {
                *reinterpret_cast<float4*>(row + xo) = o;
                if (rowM) *reinterpret_cast<float4*>(rowM + xo) = om;
            } else {   // xo == N - 2: the group wraps to the start of its row
                *reinterpret_cast<float2*>(row + xo) = make_float2(o.x, o.y);
                *reinterpret_cast<float2*>(row) = make_float2(o.z, o.w);
                if (rowM) {
                    *reinterpret_cast<float2*>(rowM + xo) = make_float2(om.x, om.y);
                    *reinterpret_cast<float2*>(rowM) = make_float2(om.z, om.w);
                }
            }
        }
    }
    const double t = block_sum<NT>(ss2, red);
    if (threadIdx.x == 0) bgStat[blockIdx.x] = make_double2(t, n);
}

// Optimiser::statImg's noise deviation and normaliseImg's scale
// (src/Optimiser.cpp:4796-4895, 4987-5001) over the P images of the call:
// stdN = mean over images of bgStddev(0, image, r) = sqrt(sum / (n - 1)) of
// the background-subtracted image (GSL's sd about a given mean), and
// scale = 1 / stdN rounded to float as the reference's RFLOAT.  One block.
__global__ __launch_bounds__(256) void k_image_scale(const double2* __restrict__ bgStat, int P,
                                                     double* __restrict__ stdN,
                                                     float* __restrict__ scale) {
    __shared__ double red[4];
    double a = 0.0;
    for (int p = threadIdx.x; p < P; p += 256) a += sqrt(bgStat[p].x / (bgStat[p].y - 1.0));
    const double tot = block_sum<256>(a, red);
    if (threadIdx.x == 0) {
        const double sN = tot / P;
        stdN[0] = sN;
        scale[0] = (float)(1.0 / sN);
    }
}

// dat[p][k] = FT_p[row(k)][iCol[k]], row = iRow wrapped to [0, N), FT_p the
// (N/2 + 1)-column half spectrum of image p.
// Grid (P, ceil(K / 256)): particles on x, the dimension with the largest limit.
// The normaliseImg scale (FFT(s y) = s FFT(y)) is applied here.
__global__ void k_img_pack(const float2* __restrict__ ft, int N, const int* __restrict__ iCol,
                           const int* __restrict__ iRow, int K, const float* __restrict__ scale,
                           float2* __restrict__ dat) {
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int p = blockIdx.x;
    const int j = iRow[k], row = j >= 0 ? j : j + N;
    const float2 f = ft[((size_t)p * N + row) * (N / 2 + 1) + iCol[k]];
    const float sc = scale[0];
    dat[(size_t)p * K + k] = make_float2(f.x * sc, f.y * sc);
}

// Initial noise spectrum (Optimiser::initSigma, src/Optimiser.cpp:5111-5209)
// from the stored half spectra F_p of the last images, nPix = N (N/2 + 1)
// pixels each.  Pass 1: per pixel, over the images of slice blockIdx.y, the
// sums of Re F, Im F and |F|^2 in FP64.
__global__ __launch_bounds__(256) void k_sigma_partial(const float2* __restrict__ ft, int P,
                                                       int nPix, double* __restrict__ part) {
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= nPix) return;
    const int G = gridDim.y, g = blockIdx.y;
    const int p0 = (int)((long long)P * g / G), p1 = (int)((long long)P * (g + 1) / G);
    double sr = 0.0, si = 0.0, sp = 0.0;
#pragma unroll 8
    for (int p = p0; p < p1; p++) {
        const float2 f = ft[(size_t)p * nPix + pix];
        sr += f.x;
        si += f.y;
        sp += (double)f.x * f.x + (double)f.y * f.y;
    }
    double* o = part + (size_t)g * 3 * nPix;
    o[pix] = sr;
    o[nPix + pix] = si;
    o[2 * (size_t)nPix + pix] = sp;
}

// Pass 2: per pixel, the G slice sums added in slice order into slice 0 (in
// place: each thread reads and writes only its own pixel).
__global__ __launch_bounds__(256) void k_sigma_fold(double* __restrict__ part, int G, int nPix) {
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= nPix) return;
    double sr = 0.0, si = 0.0, sp = 0.0;
#pragma unroll 8
    for (int g = 0; g < G; g++) {
        const double* o = part + (size_t)g * 3 * nPix;
        sr += o[pix];
        si += o[nPix + pix];
        sp += o[2 * (size_t)nPix + pix];
    }
    part[pix] = sr;
    part[nPix + pix] = si;
    part[2 * (size_t)nPix + pix] = sp;
}

// Pass 3: one block per shell u, the stored pixels (i, j) with
// rint(hypot(i, j)) == u (i the column 0 .. N/2, j the signed row), from the
// folded sums of slice 0.  The average power spectrum
// (powerSpectrum, src/Functions/Spectrum.cpp:161-190, averaged over the
// images) minus the square of the ring average of Re + Im of the average
// image (ringAverage with gsl_real_imag_sum, src/Functions/Spectrum.cpp:47-62,
// include/Complex.h:155-158), halved: sigma^2(u), and sigRcp(u) = -0.5 / sigma^2(u).
// For u < N/2 - 1 every pixel of the ring lies inside powerSpectrum's radius,
// so both averages run over the same pixels.
__global__ __launch_bounds__(256) void k_sigma_shells(const double* __restrict__ part, int N, int P,
                                                      const float* __restrict__ scale,
                                                      double* __restrict__ out, int nShell) {
    const int u = blockIdx.x, W = N / 2 + 1, nPix = N * W;
    __shared__ double red[4];
    double aP = 0.0, aA = 0.0;
    int cnt = 0;
    for (int pix = threadIdx.x; pix < nPix; pix += 256) {
        const int row = pix / W, i = pix - row * W, j = row < N / 2 ? row : row - N;
        if ((int)rint(sqrt((double)(i * i + j * j))) != u) continue;
        aP += part[2 * (size_t)nPix + pix];
        aA += part[pix] + part[nPix + pix];
        cnt++;
    }
    const double n = block_sum<256>((double)cnt, red);
    const double sumP = block_sum<256>(aP, red);
    const double sumA = block_sum<256>(aA, red);
    if (threadIdx.x == 0) {
        const double avgPs = sumP / ((double)P * n);
        const double ring = sumA / (double)P / n;
        // the spectra were taken before normaliseImg's scale: sigma^2 scales by its square
        const double sc = scale[0];
        const double s2 = 0.5 * (avgPs - ring * ring) * sc * sc;
        out[u] = s2;
        out[nShell + u] = -0.5 / s2;
    }
}

// sigRcpP[p][k] = sigRcp(iSig[k]) for every particle, as allocPreCal fills
// it for one group (src/Optimiser.cpp:8017-8022).  Grid (P, ceil(K / 256)), particles on x.
__global__ void k_sigma_apply(const double* __restrict__ rcp, const int* __restrict__ iSig, int K,
                              float* __restrict__ sig) {
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= K) return;
    sig[(size_t)blockIdx.x * K + k] = (float)rcp[iSig[k]];
}

static int fft_status(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) {
        set_error(std::string("set_particle_images: ") + what + " failed (hipFFT status " +
                  std::to_string((int)r) + ")");
        return THX_ERR_HIP;
    }
    return THX_OK;
}

void images_release(thx_ctx* c) {
    if (c->imgPlan) (void)hipfftDestroy(static_cast<hipfftHandle>(c->imgPlan));
    c->imgPlan = nullptr;
    c->imgN = c->imgP = c->imgB = 0;
    c->imgIn.release();
    c->imgReal.release();
    c->imgFT.release();
    c->imgMask.release();
    c->imgReady = false;
    c->sigPart.release();
    c->sigShell.release();
}

}  // namespace thx

using namespace thx;

int thx_ctx_set_particle_images(thx_ctx* c, const float* images, int P, int N, float bgRadius,
                                int softMask, const float* ctfP, const float* sigRcpP,
                                float* msOut) {
    clear_error();
    THX_CHECK(c && images && P > 0, THX_ERR_ARG, "set_particle_images: bad argument");
    THX_CHECK(c->K > 0, THX_ERR_STATE, "set_particle_images: pixel set not defined");
    THX_CHECK(N >= 4 && N % 2 == 0 && N <= 4096, THX_ERR_ARG,
              "set_particle_images: N must be even, 4 .. 4096");
    THX_CHECK(c->pixMaxCol <= N / 2 && c->pixMaxAbsRow < N / 2, THX_ERR_RANGE,
              "set_particle_images: the pixel set reaches beyond the images' half spectrum");
    // the sample standard deviation needs at least two background pixels
    THX_CHECK(bgRadius >= 0.f, THX_ERR_ARG, "set_particle_images: negative background radius");
    {
        const float r2 = bgRadius * bgRadius;
        long nbg = 0;
        for (int j = -N / 2; j < N / 2; j++)
            for (int i = -N / 2; i < N / 2; i++) nbg += (float)(i * i + j * j) > r2;
        THX_CHECK(nbg >= 2, THX_ERR_ARG,
                  "set_particle_images: fewer than two background pixels outside the radius");
    }
    THX_HIP(hipSetDevice(c->device));
    c->imgReady = false;
    hipStream_t st = c->stream;
    // with the soft mask, the masked images follow the unmasked ones in one
    // batch of 2 P transforms: the particle data come from the masked
    // spectra, the initial sigma from the unmasked ones
    const int B = softMask ? 2 * P : P;
    const size_t nImg = (size_t)N * N, nPix = (size_t)N * (N / 2 + 1);
    const size_t nReal = (size_t)B * nImg, nFT = (size_t)B * nPix;
    if (c->imgN != N || c->imgP != P || c->imgB != B) {
        if (c->imgPlan) (void)hipfftDestroy(static_cast<hipfftHandle>(c->imgPlan));
        c->imgPlan = nullptr;
        c->imgN = c->imgP = c->imgB = 0;
        hipfftHandle plan = nullptr;
        THX_TRY(fft_status(hipfftCreate(&plan), "plan create"));
        c->imgPlan = plan;
        int n[2] = {N, N};
        size_t work = 0;
        THX_TRY(fft_status(hipfftMakePlanMany(plan, 2, n, nullptr, 1, N * N, nullptr, 1,
                                              N * (N / 2 + 1), HIPFFT_R2C, B, &work),
                           "batched plan"));
        c->imgN = N;
        c->imgP = P;
        c->imgB = B;
    }
    hipfftHandle plan = static_cast<hipfftHandle>(c->imgPlan);
    THX_TRY(fft_status(hipfftSetStream(plan, st), "set stream"));
    THX_TRY(alloc_particles(c, P));
    THX_TRY(c->imgIn.ensure(sizeof(float) * P * nImg));
    THX_TRY(c->imgReal.ensure(sizeof(float) * nReal));
    THX_TRY(c->imgFT.ensure(sizeof(float2) * nFT));
    THX_HIP(hipMemcpyAsync(c->imgIn.p, images, sizeof(float) * P * nImg, hipMemcpyHostToDevice,
                           st));
    const size_t nk = (size_t)P * c->K;
    if (ctfP) THX_HIP(hipMemcpyAsync(c->ctf.p, ctfP, sizeof(float) * nk, hipMemcpyHostToDevice, st));
    if (sigRcpP)
        THX_HIP(hipMemcpyAsync(c->sig.p, sigRcpP, sizeof(float) * nk, hipMemcpyHostToDevice, st));
    THX_HIP(hipEventRecord(c->ev[0], st));
    // images of up to 16 K (N <= 128), 32 K (N <= 180) or, with 512-thread
    // blocks, 64 K pixels (N <= 256) are held in registers and read once;
    // others are re-read by the generic kernel
    const float r2 = bgRadius * bgRadius;
    float* masked = softMask ? c->imgReal.as<float>() + P * nImg : nullptr;
    if (softMask) {
        THX_TRY(c->imgMask.ensure(sizeof(float) * nImg));
        hipLaunchKernelGGL(k_mask_table, dim3((unsigned)((nImg + 255) / 256)), dim3(256), 0, st, N,
                           bgRadius, c->imgMask.as<float>());
        THX_HIP(hipGetLastError());
    }
    const float* mask = c->imgMask.as<const float>();
    THX_TRY(c->imgBgStat.ensure(sizeof(double2) * P));
    THX_TRY(c->imgScale.ensure(16));   // double stdN, then float scale at byte 8
    if (N % 4 == 0 && N * N <= 1024 * 16)
        hipLaunchKernelGGL((k_img_normalise_reg<16, 256>), dim3(P), dim3(256), 0, st,
                           c->imgIn.as<const float>(), N, r2, c->imgReal.as<float>(), mask,
                           masked, c->imgBgStat.as<double2>());
    else if (N % 4 == 0 && N * N <= 1024 * 32)
        hipLaunchKernelGGL((k_img_normalise_reg<32, 256>), dim3(P), dim3(256), 0, st,
                           c->imgIn.as<const float>(), N, r2, c->imgReal.as<float>(), mask,
                           masked, c->imgBgStat.as<double2>());
    else if (N % 4 == 0 && N * N <= 2048 * 32)
        hipLaunchKernelGGL((k_img_normalise_reg<32, 512>), dim3(P), dim3(512), 0, st,
                           c->imgIn.as<const float>(), N, r2, c->imgReal.as<float>(), mask,
                           masked, c->imgBgStat.as<double2>());
    else
        hipLaunchKernelGGL(k_img_normalise, dim3(P), dim3(256), 0, st, c->imgIn.as<const float>(),
                           N, r2, c->imgReal.as<float>(), mask, masked,
                           c->imgBgStat.as<double2>());
    THX_HIP(hipGetLastError());
    double* dStdN = c->imgScale.as<double>();
    float* dScale = reinterpret_cast<float*>(c->imgScale.as<char>() + 8);
    hipLaunchKernelGGL(k_image_scale, dim3(1), dim3(256), 0, st, c->imgBgStat.as<const double2>(),
                       P, dStdN, dScale);
    THX_HIP(hipGetLastError());
    THX_HIP(hipEventRecord(c->ev[1], st));
    THX_TRY(fft_status(hipfftExecR2C(plan, c->imgReal.as<hipfftReal>(), c->imgFT.as<hipfftComplex>()),
                       "forward transform"));
    THX_HIP(hipEventRecord(c->ev[2], st));
    hipLaunchKernelGGL(k_img_pack, dim3(P, (c->K + 255) / 256), dim3(256), 0, st,
                       c->imgFT.as<const float2>() + (softMask ? P * nPix : 0), N,
                       c->iCol.as<const int>(),
                       c->iRow.as<const int>(), c->K, dScale, c->dat.as<float2>());
    THX_HIP(hipGetLastError());
    THX_HIP(hipEventRecord(c->ev[3], st));
    THX_HIP(hipStreamSynchronize(st));
    c->imgReady = true;
    c->partFromImages = true;
    c->ctfReady = ctfP != nullptr;
    c->sigReady = sigRcpP != nullptr;
    c->nev = 0;   // the events now time this call
    c->evNames.clear();
    if (msOut) {
        // [0] normalise, [1] forward FFT, [2] pack
        for (int s = 0; s < 3; s++) THX_HIP(hipEventElapsedTime(msOut + s, c->ev[s], c->ev[s + 1]));
    }
    return THX_OK;
}

int thx_ctx_init_sigma(thx_ctx* c, int nShell, int apply, double* sigOut, double* sigRcpOut,
                       float* msOut) {
    clear_error();
    THX_CHECK(c, THX_ERR_ARG, "init_sigma: null context");
    THX_CHECK(c->imgReady, THX_ERR_STATE,
              "init_sigma: no particle images (call thx_ctx_set_particle_images first)");
    const int N = c->imgN, P = c->imgP;
    THX_CHECK(nShell >= 1 && nShell <= N / 2 - 1, THX_ERR_ARG,
              "init_sigma: nShell must be 1 .. N/2 - 1 (Optimiser::maxR)");
    if (apply) {
        THX_CHECK(c->partFromImages && c->P == P && c->K > 0, THX_ERR_STATE,
                  "init_sigma: the particles are not the last images");
        THX_CHECK(c->maxShell < nShell, THX_ERR_RANGE,
                  "init_sigma: the pixel set has shells beyond nShell");
    }
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int nPix = N * (N / 2 + 1), nb = (nPix + 255) / 256;
    // enough image slices for about 2048 blocks in pass 1
    const int G = std::max(1, std::min({P, 256, 2048 / nb}));
    THX_TRY(c->sigPart.ensure(sizeof(double) * 3 * (size_t)G * nPix));
    THX_TRY(c->sigShell.ensure(sizeof(double) * 2 * (size_t)nShell));
    THX_HIP(hipEventRecord(c->ev[0], st));
    hipLaunchKernelGGL(k_sigma_partial, dim3(nb, G), dim3(256), 0, st, c->imgFT.as<const float2>(), P,
                       nPix, c->sigPart.as<double>());
    THX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sigma_fold, dim3(nb), dim3(256), 0, st, c->sigPart.as<double>(), G, nPix);
    THX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sigma_shells, dim3(nShell), dim3(256), 0, st, c->sigPart.as<const double>(),
                       N, P, reinterpret_cast<const float*>(c->imgScale.as<const char>() + 8),
                       c->sigShell.as<double>(), nShell);
    THX_HIP(hipGetLastError());
    if (apply) {
        c->sigReady = true;
        hipLaunchKernelGGL(k_sigma_apply, dim3(P, (c->K + 255) / 256), dim3(256), 0, st,
                           c->sigShell.as<const double>() + nShell, c->iSig.as<const int>(), c->K,
                           c->sig.as<float>());
        THX_HIP(hipGetLastError());
    }
    THX_HIP(hipEventRecord(c->ev[1], st));
    if (sigOut)
        THX_HIP(hipMemcpyAsync(sigOut, c->sigShell.p, sizeof(double) * nShell,
                               hipMemcpyDeviceToHost, st));
    if (sigRcpOut)
        THX_HIP(hipMemcpyAsync(sigRcpOut, c->sigShell.as<double>() + nShell,
                               sizeof(double) * nShell, hipMemcpyDeviceToHost, st));
    THX_HIP(hipStreamSynchronize(st));
    c->nev = 0;
    c->evNames.clear();
    if (msOut) THX_HIP(hipEventElapsedTime(msOut, c->ev[0], c->ev[1]));
    return THX_OK;
}

int thx_ctx_get_image_stats(thx_ctx* c, double* stdN, float* scale) {
    clear_error();
    THX_CHECK(c, THX_ERR_ARG, "get_image_stats: null context");
    THX_CHECK(c->imgReady, THX_ERR_STATE,
              "get_image_stats: no particle images (call thx_ctx_set_particle_images first)");
    THX_HIP(hipSetDevice(c->device));
    THX_HIP(hipStreamSynchronize(c->stream));
    if (stdN) THX_HIP(hipMemcpy(stdN, c->imgScale.p, sizeof(double), hipMemcpyDeviceToHost));
    if (scale)
        THX_HIP(hipMemcpy(scale, c->imgScale.as<char>() + 8, sizeof(float), hipMemcpyDeviceToHost));
    return THX_OK;
}
