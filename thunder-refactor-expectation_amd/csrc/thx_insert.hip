// Reconstruction insert of the particles after an expectation, the consumer
// side of the E-step (Optimiser::reconstructRef's insert loop,
// src/Optimiser.cpp:7000-7190, with Reconstructor::insertP,
// src/Reconstructor.cpp:782-860, and Volume::addFT, src/Image/Volume.cpp:
// 334-368, 555-700; trilinear kernel, T added during the insert,
// include/Config.h:97-99).  For each particle and each of its m samples
// (class, rotation, shift): the unmasked data at the reconstruction pixel set
// (allocPreCalIdx(rU, 0), allocPreCal(mask = false)) shifted by -(t - offset),
// times the particle's CTF and the weight w / m, is added by trilinear weights
// into the class's half-complex F volume at R (i pf, j pf, 0) (conjugated and
// mirrored when x < 0), and CTF^2 w / m into the real T volume.
//
// Accumulators are FP64 and filled with device atomics (the reference adds
// with OpenMP atomics in FP32): one thread per (sample, pixel).

#include <algorithm>
#include <cmath>
#include <vector>

#include "thx_device.hpp"
#include "thx_internal.hpp"

namespace thx {

__global__ __launch_bounds__(256) void k_insert(
    const float2* __restrict__ ftN, int N, const float* __restrict__ scale,
    const int2* __restrict__ pix, int Kr, const int* __restrict__ scls,
    const double* __restrict__ quat, const double* __restrict__ trans,
    const double* __restrict__ offset, int m, const double* __restrict__ wP,
    const float* __restrict__ attr, float pixelSize, int pf, int V,
    double2* const* __restrict__ F, double* const* __restrict__ T) {
    const int s = blockIdx.x;   // sample; samples on x, the grid dimension with the largest limit
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= Kr) return;
    const int p = s / m;
    const int2 q = pix[k];
    const int i = q.x, j = q.y, row = j >= 0 ? j : j + N;
    const float2 x = ftN[((size_t)p * N + row) * (N / 2 + 1) + i];
    const float sc = scale[0];
    // translate(dst, src, -(t - offset)): phase 2 pi (i rCol + j rRow), rCol = -tx / N
    const double tx = trans[2 * (size_t)s] - (offset ? offset[2 * p] : 0.0);
    const double ty = trans[2 * (size_t)s + 1] - (offset ? offset[2 * p + 1] : 0.0);
    const float rCol = (float)(-tx) / N, rRow = (float)(-ty) / N;
    const float ph = (float)(2 * M_PI) * (i * rCol + j * rRow);
    const float cr = cosf(-ph), ci = sinf(-ph);
    const float xr = x.x * sc, xi = x.y * sc;
    const float sr = xr * cr - xi * ci, si = xr * ci + xi * cr;
    const float* a = attr + 7 * (size_t)p;
    const float lambda =
        (float)(12.2643247 / sqrt((double)a[0] * (1 + (double)a[0] * 0.978466e-6)));
    const float cv = ctf_value(i, j, pixelSize, lambda, a[1], a[2], a[3], a[4], a[5], N, a[6]);
    const float w = (float)((wP ? wP[p] : 1.0) / m);
    float vr = sr * cv * w, vi = si * cv * w;
    const float vt = cv * cv * w;
    double mm[3][2];
    quat_cols01(quat + 4 * (size_t)s, mm);
    const double c0 = (double)(i * pf), c1 = (double)(j * pf);
    float cx = (float)(mm[0][0] * c0 + mm[0][1] * c1);
    float cy = (float)(mm[1][0] * c0 + mm[1][1] * c1);
    float cz = (float)(mm[2][0] * c0 + mm[2][1] * c1);
    if (cx < 0) {   // conjHalf
        cx = -cx;
        cy = -cy;
        cz = -cz;
        vi = -vi;
    }
    const int x0 = (int)floorf(cx), y0 = (int)floorf(cy), z0 = (int)floorf(cz);
    const float xd = cx - x0, yd = cy - y0, zd = cz - z0;
    const float wx[2] = {1 - xd, xd}, wy[2] = {1 - yd, yd}, wz[2] = {1 - zd, zd};
    const int nColFT = V / 2 + 1;
    double2* f = F[scls ? scls[s] : 0];
    double* t = T[scls ? scls[s] : 0];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        int zz = z0 + c;
        zz = zz < 0 ? zz + V : zz;
#pragma unroll
        for (int b = 0; b < 2; b++) {
            int yy = y0 + b;
            yy = yy < 0 ? yy + V : yy;
            const size_t idx0 = ((size_t)zz * V + yy) * nColFT + x0;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float ww = wx[e] * wy[b] * wz[c];
                atomicAdd(&f[idx0 + e].x, (double)(vr * ww));
                atomicAdd(&f[idx0 + e].y, (double)(vi * ww));
                atomicAdd(&t[idx0 + e], (double)(vt * ww));
            }
        }
    }
}

// the pixel set of Optimiser::allocPreCalIdx(rU, 0) (src/Optimiser.cpp:
// 7942-7992): loop IMAGE_FOR_PIXEL_R_FT(rU + 1), skip (0, j < 0), keep
// i^2 + j^2 < rU^2 with rint(|(i, j)|) < rU
static std::vector<int2> recon_pixels(float rU) {
    std::vector<int2> out;
    const float r = rU + 1, rU2 = rU * rU;
    for (int j = (int)(-r); j < r; j++)
        for (int i = 0; i <= r; i++) {
            if (i == 0 && j < 0) continue;
            const float u = (float)((double)i * i + (double)j * j);
            if (u < rU2) {
                const double h = std::hypot((double)i, (double)j);
                const int v = (int)(h >= 0 ? std::floor(h + 0.5) : std::ceil(h - 0.5));
                if (v < rU) out.push_back(make_int2(i, j));
            }
        }
    return out;
}

}  // namespace thx

using namespace thx;

int thx_ctx_insert(thx_ctx* c, const int* cls, const double* quat, const double* trans,
                   const double* offset, int m, const double* w, int rU, int nCls, int clear,
                   float pixelSize, float* msOut) {
    clear_error();
    THX_CHECK(c && quat && trans && m >= 1 && nCls >= 1 && pixelSize > 0.f, THX_ERR_ARG,
              "insert: bad argument");
    THX_CHECK(c->imgReady && c->partFromImages && c->P == c->imgP, THX_ERR_STATE,
              "insert: the particles must come from thx_ctx_set_particle_images");
    THX_CHECK(c->ctfAttrOk, THX_ERR_STATE, "insert: needs the particles' CTF attributes");
    const int N = c->imgN, P = c->P, pf = c->pf, V = N * pf;
    THX_CHECK(rU >= 1 && rU <= N / 2 - 1, THX_ERR_ARG, "insert: rU must be in 1 .. N/2 - 1");
    const size_t nS = (size_t)P * m;
    THX_CHECK(nS < (1u << 31), THX_ERR_RANGE, "insert: too many samples");
    if (cls)
        for (size_t s = 0; s < nS; s++)
            THX_CHECK(cls[s] >= 0 && cls[s] < nCls, THX_ERR_ARG, "insert: class out of range");
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t nVox = (size_t)(V / 2 + 1) * V * V;
    if ((int)c->recoF.size() < nCls) {
        c->recoF.resize(nCls);
        c->recoT.resize(nCls);
    }
    if (c->recoV != V || (int)c->recoDir.size() < 4 * nCls) clear = 1;
    if (clear) {
        c->recoDir.assign(4 * (size_t)nCls, 0.0);
        c->recoV = V;
    }
    for (int k = 0; k < nCls; k++) {
        THX_TRY(c->recoF[k].ensure(sizeof(double2) * nVox));
        THX_TRY(c->recoT[k].ensure(sizeof(double) * nVox));
        if (clear) {
            THX_HIP(hipMemsetAsync(c->recoF[k].p, 0, sizeof(double2) * nVox, st));
            THX_HIP(hipMemsetAsync(c->recoT[k].p, 0, sizeof(double) * nVox, st));
        }
    }
    c->recoCls = std::max(c->recoCls, nCls);
    // insertDir: the sum of -R (t - offset, 0) per class, and the count
    for (size_t s = 0; s < nS; s++) {
        const int p = (int)(s / m), k = cls ? cls[s] : 0;
        const double* q = quat + 4 * s;
        const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
        const double R[3][2] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz)},
                                {2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz)},
                                {2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx)}};
        const double tx = trans[2 * s] - (offset ? offset[2 * p] : 0.0);
        const double ty = trans[2 * s + 1] - (offset ? offset[2 * p + 1] : 0.0);
        for (int a = 0; a < 3; a++) c->recoDir[4 * k + a] -= R[a][0] * tx + R[a][1] * ty;
        c->recoDir[4 * k + 3] += 1.0;
    }
    const std::vector<int2> pix = recon_pixels((float)rU);
    const int Kr = (int)pix.size();
    std::vector<double2*> fp(nCls);
    std::vector<double*> tp(nCls);
    for (int k = 0; k < nCls; k++) {
        fp[k] = c->recoF[k].as<double2>();
        tp[k] = c->recoT[k].as<double>();
    }
    DevBuf dPix, dCls, dQ, dT, dO, dW, dF, dTp;
    auto cleanup = [&]() {
        DevBuf* b[] = {&dPix, &dCls, &dQ, &dT, &dO, &dW, &dF, &dTp};
        for (DevBuf* x : b) x->release();
    };
    int s = THX_OK;
    if ((s = upload(dPix, pix.data(), pix.size(), st)) || (cls && (s = upload(dCls, cls, nS, st))) ||
        (s = upload(dQ, quat, 4 * nS, st)) || (s = upload(dT, trans, 2 * nS, st)) ||
        (offset && (s = upload(dO, offset, 2 * (size_t)P, st))) ||
        (w && (s = upload(dW, w, P, st))) || (s = upload(dF, fp.data(), fp.size(), st)) ||
        (s = upload(dTp, tp.data(), tp.size(), st))) {
        cleanup();
        return s;
    }
    const float* scale = reinterpret_cast<const float*>(c->imgScale.as<const char>() + 8);
    bool ok = hipEventRecord(c->ev[0], st) == hipSuccess;
    hipLaunchKernelGGL(k_insert, dim3((unsigned)nS, (Kr + 255) / 256), dim3(256), 0, st,
                       c->imgFT.as<const float2>(), N, scale, dPix.as<const int2>(), Kr,
                       cls ? dCls.as<const int>() : nullptr, dQ.as<const double>(),
                       dT.as<const double>(), offset ? dO.as<const double>() : nullptr, m,
                       w ? dW.as<const double>() : nullptr, c->ctfAttr.as<const float>(),
                       pixelSize, pf, V, dF.as<double2* const>(), dTp.as<double* const>());
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(c->ev[1], st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && msOut) ok = hipEventElapsedTime(msOut, c->ev[0], c->ev[1]) == hipSuccess;
    cleanup();
    THX_CHECK(ok, THX_ERR_HIP, "insert: kernel failed");
    c->nev = 0;
    c->evNames.clear();
    return THX_OK;
}

int thx_ctx_get_reconstruction(thx_ctx* c, int cls, double* F, double* T, double* dir,
                               double* count) {
    clear_error();
    THX_CHECK(c && cls >= 0 && cls < c->recoCls && c->recoV > 0, THX_ERR_STATE,
              "get_reconstruction: nothing inserted into this class");
    THX_HIP(hipSetDevice(c->device));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int V = c->recoV;
    const size_t nVox = (size_t)(V / 2 + 1) * V * V;
    if (F) THX_HIP(hipMemcpy(F, c->recoF[cls].p, sizeof(double2) * nVox, hipMemcpyDeviceToHost));
    if (T) THX_HIP(hipMemcpy(T, c->recoT[cls].p, sizeof(double) * nVox, hipMemcpyDeviceToHost));
    if (dir)
        for (int a = 0; a < 3; a++) dir[a] = c->recoDir[4 * cls + a];
    if (count) *count = c->recoDir[4 * cls + 3];
    return THX_OK;
}
