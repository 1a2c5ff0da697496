// Noise spectrum refreshed from the particles' best poses: the sigma^2 update
// after an expectation (Optimiser::allReduceSigma, src/Optimiser.cpp:
// 6355-6666; 3-D, best pose only (OPTIMISER_SIGMA_RANK1ST), every shell up to
// maxR (OPTIMISER_SIGMA_WHOLE_FREQUENCY), masked and unmasked residuals mixed
// (OPTIMISER_SIGMA_MASK off), translations recentred per image
// (OPTIMISER_RECENTRE_IMAGE_EACH_ITERATION), include/Config.h:165-198).
//
// Per particle, the model M = CTF . exp(-2 pi i k.t / N) . P_cls(R k) on the
// half-spectrum pixels with |k| < r (the projector's maxRadius) and zero
// beyond, M' the same with t - offset; per shell u = rint|k| < nShell the
// mean powers of X_masked - M, X_unmasked - M', M and X_masked.  Per group,
// the mean over its particles of the first two halved and of
// sqrt(|M|^2 / |X_masked|^2) give sigM, sigN and the ratio; sigma^2 =
// ratio sigM + (1 - ratio) alpha sigN with ratio = min(1, svd), alpha =
// sqrt(pi (maskRadius / (N pixelSize))^2), sigRcp = -0.5 / sigma^2.
//
// Deterministic: one block per particle, each wave a fixed set of shells with
// shuffle reductions, then one block per (group, shell) adding the particles
// in a fixed tree.

#include <algorithm>
#include <cmath>
#include <vector>

#include "thx_device.hpp"
#include "thx_internal.hpp"

namespace thx {

__device__ inline double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// pix[k] = (i, j, stored row, shell), sorted by shell; shellOff[nShell + 1].
// out[(p * nShell + u) * 3 + (0, 1, 2)] = ps(Xm - M) / 2, ps(Xn - M') / 2,
// sqrt(ps(M) / ps(Xm)).
__global__ __launch_bounds__(256) void k_sigma_refresh(
    const float2* __restrict__ ftN, const float2* __restrict__ ftM, int N,
    const float* __restrict__ scale, const int4* __restrict__ pix,
    const int* __restrict__ shellOff, int nShell, const float2* const* __restrict__ vols, int V,
    int pf, const int* __restrict__ cls, const double* __restrict__ quat,
    const double* __restrict__ trans, const double* __restrict__ offset,
    const float* __restrict__ attr, float pixelSize, int r2, double* __restrict__ out) {
    const int p = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t nPix = (size_t)N * (N / 2 + 1);
    const float2* xm = ftM + (size_t)p * nPix;
    const float2* xn = ftN + (size_t)p * nPix;
    const float2* vol = vols[cls ? cls[p] : 0];
    const float sc = scale[0];
    double m[3][2];
    quat_cols01(quat + 4 * (size_t)p, m);
    // translate(Image&, ..., r, tx, ty): phase 2 pi (i tx / N + j ty / N) in float
    const float rColM = (float)trans[2 * p] / N, rRowM = (float)trans[2 * p + 1] / N;
    const double ox = offset ? offset[2 * p] : 0.0, oy = offset ? offset[2 * p + 1] : 0.0;
    const float rColN = (float)(trans[2 * p] - ox) / N, rRowN = (float)(trans[2 * p + 1] - oy) / N;
    const float* a = attr + 7 * (size_t)p;
    const float lambda =
        (float)(12.2643247 / sqrt((double)a[0] * (1 + (double)a[0] * 0.978466e-6)));
    for (int u = wave; u < nShell; u += 4) {
        double aM = 0.0, aN = 0.0, aP = 0.0, aD = 0.0;
        for (int k = shellOff[u] + lane; k < shellOff[u + 1]; k += 64) {
            const int4 q = pix[k];
            const int i = q.x, j = q.y;
            const float2 dm = xm[(size_t)q.z * (N / 2 + 1) + i];
            const float2 dn = xn[(size_t)q.z * (N / 2 + 1) + i];
            const float xmr = dm.x * sc, xmi = dm.y * sc, xnr = dn.x * sc, xni = dn.y * sc;
            float mr = 0.f, mi = 0.f, nr = 0.f, ni = 0.f;
            if (i * i + j * j < r2) {
                const double c0 = (double)(i * pf), c1 = (double)(j * pf);
                const float x = (float)(m[0][0] * c0 + m[0][1] * c1);
                const float y = (float)(m[1][0] * c0 + m[1][1] * c1);
                const float z = (float)(m[2][0] * c0 + m[2][1] * c1);
                const float2 v = sample_trilinear(vol, V, x, y, z);
                const float cv = ctf_value(i, j, pixelSize, lambda, a[1], a[2], a[3], a[4], a[5],
                                           N, a[6]);
                const float ph = (float)(2 * M_PI) * (i * rColM + j * rRowM);
                const float cr = cosf(-ph), ci = sinf(-ph);
                mr = cv * (cr * v.x - ci * v.y);
                mi = cv * (cr * v.y + ci * v.x);
                const float ph2 = (float)(2 * M_PI) * (i * rColN + j * rRowN);
                const float cr2 = cosf(-ph2), ci2 = sinf(-ph2);
                nr = cv * (cr2 * v.x - ci2 * v.y);
                ni = cv * (cr2 * v.y + ci2 * v.x);
            }
            const double em = (double)xmr - mr, emi = (double)xmi - mi;
            const double en = (double)xnr - nr, eni = (double)xni - ni;
            aM += em * em + emi * emi;
            aN += en * en + eni * eni;
            aP += (double)mr * mr + (double)mi * mi;
            aD += (double)xmr * xmr + (double)xmi * xmi;
        }
        aM = wave_sum_d(aM);
        aN = wave_sum_d(aN);
        aP = wave_sum_d(aP);
        aD = wave_sum_d(aD);
        if (lane == 0) {
            const double cnt = (double)(shellOff[u + 1] - shellOff[u]);
            double* o = out + ((size_t)p * nShell + u) * 3;
            o[0] = aM / cnt / 2;
            o[1] = aN / cnt / 2;
            o[2] = sqrt(aP / aD);
        }
    }
}

// One block per (shell u, group g): the group's particles in a fixed order
// (order[off[g] .. off[g + 1]), the particles sorted by group; thread t adds
// every 256th, then a fixed tree), averaged and combined into sigma^2 and
// sigRcp.
__global__ __launch_bounds__(256) void k_sigma_groups(const double* __restrict__ per,
                                                      int nShell, const int* __restrict__ order,
                                                      const int* __restrict__ off, double alpha,
                                                      double* __restrict__ sig,
                                                      double* __restrict__ rcp) {
    const int u = blockIdx.x, g = blockIdx.y;
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int e = off[g] + threadIdx.x; e < off[g + 1]; e += 256) {
        const double* o = per + ((size_t)order[e] * nShell + u) * 3;
        s[0] += o[0];
        s[1] += o[1];
        s[2] += o[2];
    }
    for (int e = 0; e < 3; e++) red[e][threadIdx.x] = s[e];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int e = 0; e < 3; e++) red[e][threadIdx.x] += red[e][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = (double)(off[g + 1] - off[g]);
        const double sM = red[0][0] / n, sN = red[1][0] / n, sv = red[2][0] / n;
        const double ratio = fmin(1.0, sv);
        const double v = ratio * sM + (1.0 - ratio) * alpha * sN;
        sig[(size_t)g * nShell + u] = v;
        rcp[(size_t)g * nShell + u] = -0.5 / v;
    }
}

// sigRcpP[p][k] = sigRcp(group[p], iSig[k]) (allocPreCal, src/Optimiser.cpp:
// 8017-8022).  Grid (P, ceil(K / 256)), particles on x.
__global__ void k_sigma_apply_groups(const double* __restrict__ rcp, int nShell,
                                     const int* __restrict__ group, const int* __restrict__ iSig,
                                     int K, float* __restrict__ sig) {
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int p = blockIdx.x;
    sig[(size_t)p * K + k] = (float)rcp[(size_t)(group ? group[p] : 0) * nShell + iSig[k]];
}

}  // namespace thx

using namespace thx;

int thx_ctx_refresh_sigma(thx_ctx* c, const int* cls, const double* quat, const double* trans,
                          const double* offset, const int* group, int nGroup, int r, int nShell,
                          float maskRadius, float pixelSize, int apply, double* sigOut,
                          double* sigRcpOut, float* msOut) {
    clear_error();
    THX_CHECK(c && quat && trans && nGroup >= 1 && pixelSize > 0.f && maskRadius >= 0.f,
              THX_ERR_ARG, "refresh_sigma: bad argument");
    THX_CHECK(c->imgReady && c->partFromImages && c->P == c->imgP, THX_ERR_STATE,
              "refresh_sigma: the particles must come from thx_ctx_set_particle_images");
    THX_CHECK(c->ctfAttrOk, THX_ERR_STATE,
              "refresh_sigma: needs the particles' CTF attributes (thx_ctx_set_ctf)");
    const int N = c->imgN, P = c->P;
    THX_CHECK(nShell >= 1 && nShell <= N / 2 - 1 && r >= 1 && r <= N / 2 - 1, THX_ERR_ARG,
              "refresh_sigma: nShell and r must be in 1 .. N/2 - 1 (maxR)");
    THX_CHECK(!apply || c->maxShell < nShell, THX_ERR_RANGE,
              "refresh_sigma: the pixel set has shells beyond nShell");
    const int nCls = (int)c->vol.size();
    std::vector<int> gcount(nGroup, 0);
    for (int p = 0; p < P; p++) {
        const int k = cls ? cls[p] : 0;
        THX_CHECK(k >= 0 && k < nCls && c->vol[k].p && c->vdim[k] == N * c->pf, THX_ERR_STATE,
                  "refresh_sigma: a particle's class has no volume of the images' size");
        const int g = group ? group[p] : 0;
        THX_CHECK(g >= 0 && g < nGroup, THX_ERR_ARG, "refresh_sigma: group out of range");
        gcount[g]++;
    }
    for (int g = 0; g < nGroup; g++)
        THX_CHECK(gcount[g] > 0, THX_ERR_ARG, "refresh_sigma: a group has no particles");
    // particles sorted by group (stable), offsets per group
    std::vector<int> goff(nGroup + 1, 0), order(P);
    for (int g = 0; g < nGroup; g++) goff[g + 1] = goff[g] + gcount[g];
    {
        std::vector<int> fill(goff.begin(), goff.end() - 1);
        for (int p = 0; p < P; p++) order[fill[group ? group[p] : 0]++] = p;
    }
    // the stored half-spectrum pixels of shells 0 .. nShell - 1, by shell
    std::vector<int4> pix;
    std::vector<int> off(nShell + 1, 0);
    {
        std::vector<std::vector<int4>> by(nShell);
        for (int j = -N / 2; j < N / 2; j++)
            for (int i = 0; i <= N / 2; i++) {
                const int u = (int)std::rint(std::sqrt((double)(i * i + j * j)));
                if (u < nShell) by[u].push_back(make_int4(i, j, j >= 0 ? j : j + N, u));
            }
        for (int u = 0; u < nShell; u++) {
            off[u] = (int)pix.size();
            pix.insert(pix.end(), by[u].begin(), by[u].end());
        }
        off[nShell] = (int)pix.size();
    }
    std::vector<const float2*> vp(nCls);
    for (int k = 0; k < nCls; k++) vp[k] = c->vol[k].as<const float2>();
    const double alpha =
        std::sqrt(M_PI * std::pow((double)maskRadius / ((double)N * pixelSize), 2.0));
    THX_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf dPix, dOff, dVols, dCls, dQ, dT, dO, dG, dPer, dSig, dOrd, dGOff;
    auto cleanup = [&]() {
        DevBuf* b[] = {&dPix, &dOff, &dVols, &dCls, &dQ, &dT, &dO, &dG, &dPer, &dSig, &dOrd, &dGOff};
        for (DevBuf* x : b) x->release();
    };
    int s = THX_OK;
    if ((s = upload(dPix, pix.data(), pix.size(), st)) || (s = upload(dOff, off.data(), off.size(), st)) ||
        (s = upload(dVols, vp.data(), vp.size(), st)) || (cls && (s = upload(dCls, cls, P, st))) ||
        (s = upload(dQ, quat, 4 * (size_t)P, st)) || (s = upload(dT, trans, 2 * (size_t)P, st)) ||
        (offset && (s = upload(dO, offset, 2 * (size_t)P, st))) ||
        (group && (s = upload(dG, group, P, st))) || (s = upload(dOrd, order.data(), P, st)) ||
        (s = upload(dGOff, goff.data(), goff.size(), st)) ||
        (s = dPer.ensure(sizeof(double) * 3 * (size_t)P * nShell)) ||
        (s = dSig.ensure(sizeof(double) * 2 * (size_t)nGroup * nShell))) {
        cleanup();
        return s;
    }
    const size_t nPix = (size_t)N * (N / 2 + 1);
    const float2* ftN = c->imgFT.as<const float2>();
    const float2* ftM = c->imgB == 2 * P ? ftN + (size_t)P * nPix : ftN;   // masked, if made
    const float* scale = reinterpret_cast<const float*>(c->imgScale.as<const char>() + 8);
    double* sig = dSig.as<double>();
    double* rcp = sig + (size_t)nGroup * nShell;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    bool ok = hipEventRecord(e0, st) == hipSuccess;
    hipLaunchKernelGGL(k_sigma_refresh, dim3(P), dim3(256), 0, st, ftN, ftM, N, scale,
                       dPix.as<const int4>(), dOff.as<const int>(), nShell,
                       dVols.as<const float2* const>(), N * c->pf, c->pf,
                       cls ? dCls.as<const int>() : nullptr, dQ.as<const double>(),
                       dT.as<const double>(), offset ? dO.as<const double>() : nullptr,
                       c->ctfAttr.as<const float>(), pixelSize, r * r, dPer.as<double>());
    hipLaunchKernelGGL(k_sigma_groups, dim3(nShell, nGroup), dim3(256), 0, st,
                       dPer.as<const double>(), nShell, dOrd.as<const int>(),
                       dGOff.as<const int>(), alpha, sig, rcp);
    if (apply)
        hipLaunchKernelGGL(k_sigma_apply_groups, dim3(P, (c->K + 255) / 256), dim3(256), 0, st,
                           rcp, nShell, group ? dG.as<const int>() : nullptr,
                           c->iSig.as<const int>(), c->K, c->sig.as<float>());
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && sigOut)
        ok = hipMemcpy(sigOut, sig, sizeof(double) * nGroup * nShell, hipMemcpyDeviceToHost) ==
             hipSuccess;
    if (ok && sigRcpOut)
        ok = hipMemcpy(sigRcpOut, rcp, sizeof(double) * nGroup * nShell,
                       hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && msOut) ok = hipEventElapsedTime(msOut, e0, e1) == hipSuccess;
    cleanup();
    THX_CHECK(ok, THX_ERR_HIP, "refresh_sigma: a kernel or copy failed");
    if (apply) c->sigReady = true;
    c->nev = 0;
    c->evNames.clear();
    return THX_OK;
}
