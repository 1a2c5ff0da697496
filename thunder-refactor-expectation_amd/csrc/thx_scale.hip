// Intensity scale and noise normalisation refreshed from the particles' best
// poses, the rest of the update that follows an expectation:
//   - Optimiser::refreshScale + correctScale (src/Optimiser.cpp:5715-6023,
//     5069-5109): per group the ratio sum Re(X M) c / sum Re(M M) c^2 of the
//     unmasked data X against the shifted best-pose slice M and the CTF c
//     (scaleDataVSPrior, :10021-10063), clamped about the median; applied by
//     dividing the images by their group's scale and sigma^2 by its square,
//     or (the initial call) by scaling the class volumes;
//   - Optimiser::normCorrection (:6161-6351): per particle the power of the
//     masked data minus the CTF-weighted shifted slice over
//     rL <= |k| < rNorm; every image multiplied by sqrt(median / power).
// Switches as include/Config.h sets them: OPTIMISER_SCALE_MASK off,
// OPTIMISER_NORM_MASK on, OPTIMISER_REFRESH_SCALE_SPECTRUM off,
// OPTIMISER_REFRESH_SCALE_ZERO_FREQ_NO_COORD on, recentred images.
//
// Deterministic: one block per particle with fixed-order reductions, the
// per-group sums over particles sorted by group in a fixed order, and the
// median as an exact radix selection on the device.

#include <algorithm>
#include <cmath>
#include <vector>

#include "thx_device.hpp"
#include "thx_internal.hpp"

namespace thx {

namespace {

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum of v over a 256-thread block in a fixed order; the result in thread 0.
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return s;
}

// The best-pose model of particle p at pixel (i, j): CTF c and the slice
// P_cls(R (i pf, j pf, 0)) (zero when i^2 + j^2 >= r2), shifted by
// exp(-2 pi i (i rCol + j rRow)) when shift is set.
__device__ inline void best_model(const float2* vol, int V, int pf, const double (&m)[3][2],
                                  int i, int j, int r2, bool shift, float rCol, float rRow,
                                  const float* a, float lambda, float pixelSize, int N,
                                  float& mr, float& mi, float& cv) {
    cv = ctf_value(i, j, pixelSize, lambda, a[1], a[2], a[3], a[4], a[5], N, a[6]);
    mr = 0.f;
    mi = 0.f;
    if (i * i + j * j >= r2) return;
    const double c0 = (double)(i * pf), c1 = (double)(j * pf);
    const float x = (float)(m[0][0] * c0 + m[0][1] * c1);
    const float y = (float)(m[1][0] * c0 + m[1][1] * c1);
    const float z = (float)(m[2][0] * c0 + m[2][1] * c1);
    const float2 v = sample_trilinear(vol, V, x, y, z);
    if (!shift) {
        mr = v.x;
        mi = v.y;
        return;
    }
    const float ph = (float)(2 * M_PI) * (i * rCol + j * rRow);
    const float cr = cosf(-ph), ci = sinf(-ph);
    mr = cr * v.x - ci * v.y;
    mi = cr * v.y + ci * v.x;
}

__device__ inline float electron_lambda(const float* a) {
    return (float)(12.2643247 / sqrt((double)a[0] * (1 + (double)a[0] * 0.978466e-6)));
}

// refreshScale's per-particle sums over the pixels pix[0..nPix) (i, j, stored
// row): out[2p] = sum Re(X M) c, out[2p + 1] = sum Re(M M) c^2, X the
// unmasked data times the normalisation scale.  With coord == 0 the model is
// class 0's unshifted slice.
__global__ __launch_bounds__(256) void k_scale_terms(
    const float2* __restrict__ ftN, int N, const float* __restrict__ scale,
    const int4* __restrict__ pix, int nPix, const float2* const* __restrict__ vols, int V, int pf,
    int coord, const int* __restrict__ cls, const double* __restrict__ quat,
    const double* __restrict__ trans, const double* __restrict__ offset,
    const float* __restrict__ attr, float pixelSize, int r2, double* __restrict__ out) {
    __shared__ double red[4];
    const int p = blockIdx.x;
    const float2* x = ftN + (size_t)p * N * (N / 2 + 1);
    const float2* vol = vols[coord && cls ? cls[p] : 0];
    const float sc = scale[0];
    double m[3][2];
    quat_cols01(quat + 4 * (size_t)p, m);
    float rCol = 0.f, rRow = 0.f;
    if (coord) {
        const double tx = trans[2 * p] - (offset ? offset[2 * p] : 0.0);
        const double ty = trans[2 * p + 1] - (offset ? offset[2 * p + 1] : 0.0);
        rCol = (float)tx / N;
        rRow = (float)ty / N;
    }
    const float* a = attr + 7 * (size_t)p;
    const float lambda = electron_lambda(a);
    double sXA = 0.0, sAA = 0.0;
    for (int k = threadIdx.x; k < nPix; k += 256) {
        const int4 q = pix[k];
        float mr, mi, cv;
        best_model(vol, V, pf, m, q.x, q.y, r2, coord != 0, rCol, rRow, a, lambda, pixelSize, N,
                   mr, mi, cv);
        const float2 d = x[(size_t)q.z * (N / 2 + 1) + q.x];
        const float xr = d.x * sc, xi = d.y * sc;
        sXA += (double)((xr * mr - xi * mi) * cv);
        sAA += (double)((mr * mr - mi * mi) * (cv * cv));
    }
    sXA = block_sum(sXA, red);
    sAA = block_sum(sAA, red);
    if (threadIdx.x == 0) {
        out[2 * (size_t)p] = sXA;
        out[2 * (size_t)p + 1] = sAA;
    }
}

// normCorrection's per-particle power sum |X_masked - c M|^2 over pix.
__global__ __launch_bounds__(256) void k_norm_terms(
    const float2* __restrict__ ftM, int N, const float* __restrict__ scale,
    const int4* __restrict__ pix, int nPix, const float2* const* __restrict__ vols, int V, int pf,
    const int* __restrict__ cls, const double* __restrict__ quat,
    const double* __restrict__ trans, const float* __restrict__ attr, float pixelSize, int r2,
    double* __restrict__ norm) {
    __shared__ double red[4];
    const int p = blockIdx.x;
    const float2* x = ftM + (size_t)p * N * (N / 2 + 1);
    const float2* vol = vols[cls ? cls[p] : 0];
    const float sc = scale[0];
    double m[3][2];
    quat_cols01(quat + 4 * (size_t)p, m);
    const float rCol = (float)trans[2 * p] / N, rRow = (float)trans[2 * p + 1] / N;
    const float* a = attr + 7 * (size_t)p;
    const float lambda = electron_lambda(a);
    double s = 0.0;
    for (int k = threadIdx.x; k < nPix; k += 256) {
        const int4 q = pix[k];
        float mr, mi, cv;
        best_model(vol, V, pf, m, q.x, q.y, r2, true, rCol, rRow, a, lambda, pixelSize, N, mr,
                   mi, cv);
        const float2 d = x[(size_t)q.z * (N / 2 + 1) + q.x];
        const float er = d.x * sc - mr * cv, ei = d.y * sc - mi * cv;
        s += (double)er * er + (double)ei * ei;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) norm[p] = s;
}

// Per group g the sums of the pairs per[2p], per[2p + 1] of its particles,
// taken in the order order[off[g] .. off[g + 1]) (particles sorted by group),
// thread t adding every 256th, then a fixed tree.  One block per group.
__global__ __launch_bounds__(256) void k_group_pair_sums(const double* __restrict__ per,
                                                         const int* __restrict__ order,
                                                         const int* __restrict__ off,
                                                         double* __restrict__ gsum) {
    __shared__ double red[4];
    const int g = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (int e = off[g] + threadIdx.x; e < off[g + 1]; e += 256) {
        const int p = order[e];
        a += per[2 * (size_t)p];
        b += per[2 * (size_t)p + 1];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    if (threadIdx.x == 0) {
        gsum[2 * g] = a;
        gsum[2 * g + 1] = b;
    }
}

// The k-th smallest (0-based) of n non-negative doubles, exactly: their bit
// patterns order as unsigned integers, so an MSB-first radix selection with
// 8-bit digits (eight histogram passes over the values) finds its pattern.
// One block of 1024 threads; the result in sel[0].
__device__ unsigned long long radix_select(const double* __restrict__ v, int n, int k,
                                           unsigned int* hist, unsigned long long* sh) {
    unsigned long long prefix = 0, mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(v[e]);
            if ((bits & mask) == prefix) atomicAdd(&hist[(bits >> shift) & 255], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int b = 0;
            unsigned int below = 0;
            while (below + hist[b] <= (unsigned int)k) below += hist[b++];
            sh[0] = (unsigned long long)b;
            sh[1] = below;
        }
        __syncthreads();
        const unsigned long long b = sh[0];
        k -= (int)sh[1];
        prefix |= b << shift;
        mask |= 255ull << shift;
        __syncthreads();
    }
    return prefix;
}

// median(vec, n) of the powers (src/Functions/Functions.cpp:227-233: GSL's
// 0.5 quantile of the sorted values) rounded to float as the reference's
// RFLOAT m, then factor[p] = sqrt(m / norm[p]).  med[0] receives m.
__global__ __launch_bounds__(1024) void k_norm_median(const double* __restrict__ norm, int n,
                                                      double* __restrict__ med) {
    __shared__ unsigned int hist[256];
    __shared__ unsigned long long sh[2];
    const int lhs = (n - 1) / 2;
    const double a = __longlong_as_double((long long)radix_select(norm, n, lhs, hist, sh));
    double m = a;
    if ((n & 1) == 0) {
        const double b =
            __longlong_as_double((long long)radix_select(norm, n, lhs + 1, hist, sh));
        m = 0.5 * a + 0.5 * b;
    }
    if (threadIdx.x == 0) med[0] = (double)(float)m;
}

// Per-particle factors of the images and of sigRcp.
//   mode 0 (normCorrection): fImg = sqrt(med / norm[p]), no sigma factor;
//   mode 1 (correctScale):   fImg = 1 / s, fSig = s^2 with s = scale[group[p]].
// The image factor is rounded to float, as the reference multiplies RFLOAT
// pixels by it.
__global__ void k_particle_factors(int P, int mode, const double* __restrict__ norm,
                                   const double* __restrict__ med,
                                   const double* __restrict__ scale,
                                   const int* __restrict__ group, float* __restrict__ fImg,
                                   float* __restrict__ fSig) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (mode == 0) {
        fImg[p] = (float)sqrt(med[0] / norm[p]);
    } else {
        const double s = scale[group ? group[p] : 0];
        fImg[p] = (float)(1.0 / s);
        fSig[p] = (float)(s * s);
    }
}

// Scale particle p's stored spectra (nSpec of them: unmasked, then masked at
// +P), its packed data and, with fSig, its sigRcp row.  Grid (P, chunks of
// 256 x 4 elements), particles on x.
__global__ __launch_bounds__(256) void k_scale_particles(float2* __restrict__ ft, int P,
                                                         int nSpec, size_t nPix,
                                                         float2* __restrict__ dat,
                                                         float* __restrict__ sig, int K,
                                                         const float* __restrict__ fImg,
                                                         const float* __restrict__ fSig) {
    const int p = blockIdx.x;
    const float f = fImg[p];
    const size_t base = ((size_t)blockIdx.y * 256 + threadIdx.x) * 4;
    for (int s = 0; s < nSpec; s++) {
        float2* x = ft + ((size_t)s * P + p) * nPix;
        for (int e = 0; e < 4; e++) {
            const size_t i = base + e;
            if (i < nPix) {
                const float2 v = x[i];
                x[i] = make_float2(v.x * f, v.y * f);
            }
        }
    }
    for (int e = 0; e < 4; e++) {
        const size_t i = base + e;
        if (i < (size_t)K) {
            const float2 v = dat[(size_t)p * K + i];
            dat[(size_t)p * K + i] = make_float2(v.x * f, v.y * f);
            if (fSig) sig[(size_t)p * K + i] *= fSig[p];
        }
    }
}

// vol[i] *= s (SCALE_FT on a class reference before its projectee is built;
// the projectee is linear in it).
__global__ void k_scale_volume(float2* __restrict__ vol, size_t n, float s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = vol[i];
    vol[i] = make_float2(v.x * s, v.y * s);
}

// Stored half-spectrum pixels (i = 0 .. N/2, j = -N/2 .. N/2 - 1, both halves
// of the i = 0 column) that pass keep(i, j).
template <class F>
std::vector<int4> spectrum_pixels(int N, F keep) {
    std::vector<int4> out;
    for (int j = -N / 2; j < N / 2; j++)
        for (int i = 0; i <= N / 2; i++)
            if (keep(i, j)) out.push_back(make_int4(i, j, j >= 0 ? j : j + N, 0));
    return out;
}

// Common checks of the best-pose entry points; fills the volume pointers.
int best_pose_checks(thx_ctx* c, const char* name, const int* cls, int useCls,
                     std::vector<const float2*>& vp) {
    THX_CHECK(c->imgReady && c->partFromImages && c->P == c->imgP, THX_ERR_STATE,
              std::string(name) + ": the particles must come from thx_ctx_set_particle_images");
    THX_CHECK(c->ctfAttrOk, THX_ERR_STATE,
              std::string(name) + ": needs the particles' CTF attributes (thx_ctx_set_ctf)");
    const int nCls = (int)c->vol.size(), N = c->imgN;
    for (int p = 0; p < c->P; p++) {
        const int k = useCls && cls ? cls[p] : 0;
        THX_CHECK(k >= 0 && k < nCls && c->vol[k].p && c->vdim[k] == N * c->pf, THX_ERR_STATE,
                  std::string(name) + ": a particle's class has no volume of the images' size");
        if (!useCls) break;
    }
    vp.resize(nCls);
    for (int k = 0; k < nCls; k++) vp[k] = c->vol[k].as<const float2>();
    return THX_OK;
}

}  // namespace

}  // namespace thx

using namespace thx;

int thx_ctx_refresh_scale(thx_ctx* c, int coord, const int* cls, const double* quat,
                          const double* trans, const double* offset, const int* group,
                          int nGroup, int r, int rS, float rL, float pixelSize, int apply,
                          double* scaleOut, float* msOut) {
    clear_error();
    THX_CHECK(c && nGroup >= 1 && pixelSize > 0.f && apply >= 0 && apply <= 2, THX_ERR_ARG,
              "refresh_scale: bad argument");
    THX_CHECK(!coord || (quat && trans), THX_ERR_ARG,
              "refresh_scale: best poses need quaternions and translations");
    std::vector<const float2*> vp;
    THX_TRY(best_pose_checks(c, "refresh_scale", cls, coord, vp));
    const int N = c->imgN, P = c->P;
    THX_CHECK(r >= 1 && r <= N / 2 - 1, THX_ERR_ARG, "refresh_scale: r must be in 1 .. N/2 - 1");
    if (!coord) {   // OPTIMISER_REFRESH_SCALE_ZERO_FREQ_NO_COORD
        rS = 1;
        rL = 0.f;
    }
    rS = std::min(rS, r);   // "_rS is Larger than _r, Set _rS to _r"
    THX_CHECK(rS >= 1 && rL >= 0.f && rL < (float)rS, THX_ERR_ARG,
              "refresh_scale: needs 0 <= rL < rS");
    if (group)
        for (int p = 0; p < P; p++)
            THX_CHECK(group[p] >= 0 && group[p] < nGroup, THX_ERR_ARG,
                      "refresh_scale: group out of range");
    // particles sorted by group (stable), offsets per group
    const int nG = group ? nGroup : 1;
    std::vector<int> offs(nG + 1, 0), order(P);
    for (int p = 0; p < P; p++) offs[(group ? group[p] : 0) + 1]++;
    for (int g = 0; g < nG; g++) offs[g + 1] += offs[g];
    {
        std::vector<int> fill(offs.begin(), offs.end() - 1);
        for (int p = 0; p < P; p++) order[fill[group ? group[p] : 0]++] = p;
    }
    for (int g = 0; g < nG; g++)
        THX_CHECK(offs[g + 1] > offs[g], THX_ERR_ARG, "refresh_scale: a group has no particles");
    const double rS2 = (double)rS * rS, rL2 = (double)rL * (double)rL;
    const std::vector<int4> pix = spectrum_pixels(N, [&](int i, int j) {
        const double u = (double)i * i + (double)j * j;
        const double v = std::rint(std::hypot((double)i, (double)j));
        return u < rS2 && u >= rL2 && v < rS && v >= rL;
    });
    std::vector<double> q0;
    if (!quat) q0.assign(4 * (size_t)P, 0.0);
    if (!quat)
        for (int p = 0; p < P; p++) q0[4 * (size_t)p] = 1.0;
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dPix, dVols, dCls, dQ, dT, dO, dG, dOrd, dOff, dPer, dGs, dScale, dF;
    auto cleanup = [&]() {
        DevBuf* b[] = {&dPix, &dVols, &dCls, &dQ, &dT, &dO, &dG, &dOrd, &dOff, &dPer, &dGs,
                       &dScale, &dF};
        for (DevBuf* x : b) x->release();
    };
    int s = THX_OK;
    if ((s = upload(dPix, pix.data(), pix.size(), st)) ||
        (s = upload(dVols, vp.data(), vp.size(), st)) ||
        (coord && cls && (s = upload(dCls, cls, P, st))) ||
        (s = upload(dQ, quat ? quat : q0.data(), 4 * (size_t)P, st)) ||
        (coord && (s = upload(dT, trans, 2 * (size_t)P, st))) ||
        (coord && offset && (s = upload(dO, offset, 2 * (size_t)P, st))) ||
        (group && (s = upload(dG, group, P, st))) || (s = upload(dOrd, order.data(), P, st)) ||
        (s = upload(dOff, offs.data(), offs.size(), st)) ||
        (s = dPer.ensure(sizeof(double) * 2 * (size_t)P)) ||
        (s = dGs.ensure(sizeof(double) * 2 * (size_t)nG)) ||
        (s = dF.ensure(sizeof(float) * 2 * (size_t)P))) {
        cleanup();
        return s;
    }
    const float* scl = reinterpret_cast<const float*>(c->imgScale.as<const char>() + 8);
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1], e2 = c->ev[2];
    bool ok = hipEventRecord(e0, st) == hipSuccess;
    hipLaunchKernelGGL(k_scale_terms, dim3(P), dim3(256), 0, st, c->imgFT.as<const float2>(), N,
                       scl, dPix.as<const int4>(), (int)pix.size(),
                       dVols.as<const float2* const>(), N * c->pf, c->pf, coord,
                       coord && cls ? dCls.as<const int>() : nullptr, dQ.as<const double>(),
                       coord ? dT.as<const double>() : nullptr,
                       coord && offset ? dO.as<const double>() : nullptr,
                       c->ctfAttr.as<const float>(), pixelSize, r * r, dPer.as<double>());
    hipLaunchKernelGGL(k_group_pair_sums, dim3(nG), dim3(256), 0, st, dPer.as<const double>(),
                       dOrd.as<const int>(), dOff.as<const int>(), dGs.as<double>());
    std::vector<double> gs(2 * (size_t)nG);
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
         hipMemcpyAsync(gs.data(), dGs.p, sizeof(double) * gs.size(), hipMemcpyDeviceToHost,
                        st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    float ms = 0.f;
    if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (!ok) {
        cleanup();
        THX_CHECK(false, THX_ERR_HIP, "refresh_scale: a kernel or copy failed");
    }
    // per group the ratio of sums, then the clamp about the median (nGroup
    // numbers: host arithmetic, as the reference's master does it)
    std::vector<double> scale(nGroup);
    for (int g = 0; g < nGroup; g++) {
        const int h = group ? g : 0;
        scale[g] = gs[2 * h] / gs[2 * h + 1];
    }
    std::vector<double> sorted(scale);
    std::sort(sorted.begin(), sorted.end());
    const double idx = 0.5 * (nGroup - 1);
    const int lhs = (int)std::floor(idx);
    const double d = idx - lhs;
    const float med = (float)(lhs == nGroup - 1 ? sorted[lhs]
                                                : (1 - d) * sorted[lhs] + d * sorted[lhs + 1]);
    double mean = 0.0;
    for (int g = 0; g < nGroup; g++) {
        if (std::fabs(scale[g]) > std::fabs(med * 5))
            scale[g] = med * 5;
        else if (std::fabs(scale[g]) < std::fabs(med / 5))
            scale[g] = med / 5;
        mean += scale[g];
    }
    mean /= nGroup;
    if (scaleOut) std::copy(scale.begin(), scale.end(), scaleOut);
    if (!(mean >= 0.0)) {
        cleanup();
        THX_CHECK(false, THX_ERR_RANGE,
                  "refresh_scale: the average intensity scale is negative (or not a number)");
    }
    if (apply == 1) {   // correctScale(init = false): images / scale, sigma^2 / scale^2
        const size_t nPix = (size_t)N * (N / 2 + 1);
        const int nSpec = c->imgB == 2 * P ? 2 : 1;
        const size_t nEl = std::max(nPix, (size_t)c->K);
        ok = upload(dScale, scale.data(), scale.size(), st) == THX_OK &&
             hipEventRecord(e1, st) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(k_particle_factors, dim3((P + 255) / 256), dim3(256), 0, st, P, 1,
                               nullptr, nullptr, dScale.as<const double>(),
                               group ? dG.as<const int>() : nullptr, dF.as<float>(),
                               dF.as<float>() + P);
            hipLaunchKernelGGL(k_scale_particles, dim3(P, (unsigned)((nEl + 1023) / 1024)),
                               dim3(256), 0, st, c->imgFT.as<float2>(), P, nSpec, nPix,
                               c->dat.as<float2>(), c->sigReady ? c->sig.as<float>() : nullptr,
                               c->K, dF.as<const float>(),
                               c->sigReady ? dF.as<const float>() + P : nullptr);
            ok = hipGetLastError() == hipSuccess && hipEventRecord(e2, st) == hipSuccess &&
                 hipStreamSynchronize(st) == hipSuccess;
            float ms2 = 0.f;
            if (ok) ok = hipEventElapsedTime(&ms2, e1, e2) == hipSuccess;
            ms += ms2;
        }
    } else if (apply == 2) {   // correctScale(init = true): every class reference x scale
        const float sv = (float)scale[group ? group[0] : 0];
        ok = hipEventRecord(e1, st) == hipSuccess;
        for (int k = 0; ok && k < (int)c->vol.size(); k++) {
            if (!c->vol[k].p) continue;
            const int Vk = c->vdim[k];
            const size_t n = (size_t)(Vk / 2 + 1) * Vk * Vk;
            hipLaunchKernelGGL(k_scale_volume, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                               st, c->vol[k].as<float2>(), n, sv);
            invalidate_vol_lo(c, k);
        }
        ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e2, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        float ms2 = 0.f;
        if (ok) ok = hipEventElapsedTime(&ms2, e1, e2) == hipSuccess;
        ms += ms2;
    }
    cleanup();
    THX_CHECK(ok, THX_ERR_HIP, "refresh_scale: applying the scale failed");
    if (msOut) *msOut = ms;
    c->nev = 0;
    c->evNames.clear();
    return THX_OK;
}

int thx_ctx_norm_correction(thx_ctx* c, const int* cls, const double* quat,
                            const double* trans, int r, float rL, float rNorm, float pixelSize,
                            int apply, double* normOut, double* medianOut, float* msOut) {
    clear_error();
    THX_CHECK(c && quat && trans && pixelSize > 0.f, THX_ERR_ARG,
              "norm_correction: bad argument");
    std::vector<const float2*> vp;
    THX_TRY(best_pose_checks(c, "norm_correction", cls, 1, vp));
    const int N = c->imgN, P = c->P;
    THX_CHECK(r >= 1 && r <= N / 2 - 1, THX_ERR_ARG,
              "norm_correction: r must be in 1 .. N/2 - 1");
    rNorm = std::min(rNorm, (float)r);   // rNorm = min(_r, resolutionP(0.75))
    THX_CHECK(rL >= 0.f && rL < rNorm, THX_ERR_ARG, "norm_correction: needs 0 <= rL < rNorm");
    const double rN2 = (double)rNorm * (double)rNorm, rL2 = (double)rL * (double)rL;
    const std::vector<int4> pix = spectrum_pixels(N, [&](int i, int j) {
        const double u = (double)i * i + (double)j * j;
        return u >= rL2 && u < rN2;
    });
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dPix, dVols, dCls, dQ, dT, dNorm, dF;
    auto cleanup = [&]() {
        DevBuf* b[] = {&dPix, &dVols, &dCls, &dQ, &dT, &dNorm, &dF};
        for (DevBuf* x : b) x->release();
    };
    int s = THX_OK;
    if ((s = upload(dPix, pix.data(), pix.size(), st)) ||
        (s = upload(dVols, vp.data(), vp.size(), st)) || (cls && (s = upload(dCls, cls, P, st))) ||
        (s = upload(dQ, quat, 4 * (size_t)P, st)) || (s = upload(dT, trans, 2 * (size_t)P, st)) ||
        (s = dNorm.ensure(sizeof(double) * ((size_t)P + 1))) ||
        (s = dF.ensure(sizeof(float) * (size_t)P))) {
        cleanup();
        return s;
    }
    const size_t nPix = (size_t)N * (N / 2 + 1);
    const float2* ftN = c->imgFT.as<const float2>();
    const float2* ftM = c->imgB == 2 * P ? ftN + (size_t)P * nPix : ftN;   // masked, if made
    const float* scl = reinterpret_cast<const float*>(c->imgScale.as<const char>() + 8);
    double* norm = dNorm.as<double>();
    double* med = norm + P;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    bool ok = hipEventRecord(e0, st) == hipSuccess;
    hipLaunchKernelGGL(k_norm_terms, dim3(P), dim3(256), 0, st, ftM, N, scl,
                       dPix.as<const int4>(), (int)pix.size(), dVols.as<const float2* const>(),
                       N * c->pf, c->pf, cls ? dCls.as<const int>() : nullptr,
                       dQ.as<const double>(), dT.as<const double>(),
                       c->ctfAttr.as<const float>(), pixelSize, r * r, norm);
    hipLaunchKernelGGL(k_norm_median, dim3(1), dim3(1024), 0, st, norm, P, med);
    if (apply) {
        hipLaunchKernelGGL(k_particle_factors, dim3((P + 255) / 256), dim3(256), 0, st, P, 0,
                           norm, med, nullptr, nullptr, dF.as<float>(), nullptr);
        const int nSpec = c->imgB == 2 * P ? 2 : 1;
        const size_t nEl = std::max(nPix, (size_t)c->K);
        hipLaunchKernelGGL(k_scale_particles, dim3(P, (unsigned)((nEl + 1023) / 1024)),
                           dim3(256), 0, st, c->imgFT.as<float2>(), P, nSpec, nPix,
                           c->dat.as<float2>(), nullptr, c->K, dF.as<const float>(), nullptr);
    }
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && normOut)
        ok = hipMemcpy(normOut, norm, sizeof(double) * P, hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && medianOut)
        ok = hipMemcpy(medianOut, med, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && msOut) ok = hipEventElapsedTime(msOut, e0, e1) == hipSuccess;
    cleanup();
    THX_CHECK(ok, THX_ERR_HIP, "norm_correction: a kernel or copy failed");
    c->nev = 0;
    c->evNames.clear();
    return THX_OK;
}
