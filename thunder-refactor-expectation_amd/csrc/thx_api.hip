// C ABI of the expectation step: context management, the reference-shaped
// entry points and the synthetic workload generator (host side).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "thx_internal.hpp"

namespace thx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

int DevBuf::ensure(size_t need) {
    if (need <= bytes && p) return THX_OK;
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        bytes = 0;
        if (e != hipSuccess) {
            set_error(std::string("hipFree failed: ") + hipGetErrorString(e));
            return THX_ERR_HIP;
        }
    }
    if (need == 0) return THX_OK;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        set_error("hipMalloc of " + std::to_string(need) + " bytes failed: " + hipGetErrorString(e));
        return THX_ERR_OOM;
    }
    bytes = need;
    return THX_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int HostBuf::ensure(size_t need) {
    if (need <= bytes && p) return THX_OK;
    release();
    if (need == 0) return THX_OK;
    hipError_t e = hipHostMalloc(&p, need, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        set_error("hipHostMalloc of " + std::to_string(need) + " bytes failed: " + hipGetErrorString(e));
        return THX_ERR_OOM;
    }
    bytes = need;
    return THX_OK;
}

void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

// Translation chunking of a scan: rows per particle TC in {32, 64, 128}.
// The plain kernel minimises the padded count ceil(T / TC) * TC (ties to the
// larger TC); the 16-bit kernel prefers TC = 32 at equal padding, which lets
// four particles share each phase-table row the block loads.  THX_TC16 in
// the environment overrides the 16-bit choice (32, 64 or 128).
static void choose_tc(int T, bool preferSmall, int& TC, int& nTc) {
    const int order[3] = {128, 64, 32};
    int best = -1, bestPad = 0;
    for (int i = 0; i < 3; i++) {
        const int tc = preferSmall ? order[2 - i] : order[i];
        const int pad = ((T + tc - 1) / tc) * tc;
        if (best < 0 || pad < bestPad) { best = tc; bestPad = pad; }
    }
    TC = best;
    nTc = (T + best - 1) / best;
}

static int choose_scan_shape(thx_ctx* c, int precision) {
    const bool lo = precision != THX_FP32;
    choose_tc(c->T, lo, c->TC, c->nTc);
    if (lo) {
        if (const char* e = std::getenv("THX_TC16")) {
            const int tc = std::atoi(e);
            if (tc == 32 || tc == 64 || tc == 128) {
                c->TC = tc;
                c->nTc = (c->T + tc - 1) / tc;
            }
        }
    }
    c->Tpad = c->TC * c->nTc;
    c->NB = score_tile_cols(precision);
    c->nRb = c->Rpad / c->NB;
    THX_CHECK(c->Tpad <= c->TpadS, THX_ERR_STATE, "scan shape: translation padding too large");
    THX_CHECK((long long)c->nTc * c->nRb <= kMaxTiles, THX_ERR_RANGE,
              "scan shape: too many rotation/translation tiles per particle");
    return THX_OK;
}

static int set_device(thx_ctx* c) {
    THX_HIP(hipSetDevice(c->device));
    return THX_OK;
}

static int ensure_grid_buffers(thx_ctx* c) {
    THX_TRY(c->rotP.ensure(sizeof(float2) * (size_t)c->Rpad * c->Kp));
    THX_TRY(c->normP.ensure(sizeof(float) * (size_t)c->Rpad * c->Kp));
    THX_TRY(c->S.ensure(sizeof(float2) * (size_t)c->TpadS * c->Kp));
    return THX_OK;
}

static int ensure_scan_buffers(thx_ctx* c, int cls) {
    const size_t P = (size_t)c->P;
    const size_t nt = (size_t)c->nTc * c->nRb;
    THX_TRY(c->Cterm.ensure(sizeof(float) * (size_t)c->Ppad * c->Rpad));
    THX_TRY(c->partMax.ensure(sizeof(float) * P * nt));
    THX_TRY(c->partWC.ensure(sizeof(float) * P * nt));
    THX_TRY(c->partArg.ensure(sizeof(int) * P * nt));
    THX_TRY(c->partWR.ensure(sizeof(float) * P * c->nTc * c->Rpad));
    THX_TRY(c->partWT.ensure(sizeof(float) * P * c->nRb * c->Tpad));
    if ((int)c->res.size() <= cls) c->res.resize(cls + 1);
    ClassResult& r = c->res[cls];
    THX_TRY(r.wC.ensure(sizeof(double) * P));
    THX_TRY(r.wR.ensure(sizeof(double) * P * c->R));
    THX_TRY(r.wT.ensure(sizeof(double) * P * c->T));
    THX_TRY(r.base.ensure(sizeof(double) * P));
    THX_TRY(r.bestR.ensure(sizeof(int) * P));
    THX_TRY(r.bestT.ensure(sizeof(int) * P));
    THX_TRY(r.bestScore.ensure(sizeof(double) * P));
    return THX_OK;
}

static int record(thx_ctx* c, const char* name) {
    if (c->nev >= 16) return THX_OK;
    THX_HIP(hipEventRecord(c->ev[c->nev], c->stream));
    c->evNames.push_back(name);
    c->nev++;
    return THX_OK;
}

// Scoring of the particles against the current rotP / normP / S; the
// projection (or the caller's upload) has already filled rotP and normP.
static int run_scan(thx_ctx* c, int cls, int precision) {
    THX_TRY(choose_scan_shape(c, precision));
    THX_TRY(ensure_scan_buffers(c, cls));
    if (precision == THX_FP16 || precision == THX_FP8) {
        THX_TRY(c->shellScale.ensure(sizeof(float) * (c->maxShell + 2)));   // + global factor
        const void* was = c->shellMax.p;
        THX_TRY(c->shellMax.ensure(sizeof(float) * 2 * (c->maxShell + 1)));
        if (c->shellMax.p != was) c->shellMaxPOk = false;
        THX_TRY(launch_shell_scale(c, precision));
        THX_TRY(record(c, "shell_scale"));
    }
    THX_TRY(launch_prep_particles(c, precision));
    THX_TRY(record(c, "prep"));
    if (precision != THX_FP32) {
        const void* was = c->rotLo.p;
        THX_TRY(c->rotLo.ensure(sizeof(uint32_t) * (size_t)c->Rpad * c->Kp));
        if (c->rotLo.p != was) c->rotLoPrec = -1;
        // the BF16 table depends on the projections only (no particle-derived
        // scales): kept when it already holds the current ones
        if (precision != THX_BF16 || c->rotLoPrec != THX_BF16) THX_TRY(launch_convert_lo(c, precision));
        THX_TRY(record(c, "convert"));
    }
    THX_TRY(launch_norm_gemm(c, precision));
    THX_TRY(record(c, "norm_gemm"));
    THX_TRY(launch_score(c, precision));
    THX_TRY(record(c, "score"));
    THX_TRY(launch_merge(c, cls));
    THX_TRY(record(c, "merge"));
    c->res[cls].valid = true;
    return THX_OK;
}


// Particles sorted by the pose of one rotation each (quat + 4 off[p]): by
// class, then along a Morton curve over (x, y, z) of the quaternion taken with
// w >= 0, so neighbours in the list project nearby slices of the same volume.
// 30-bit Morton code of (x, y, z) of a unit quaternion taken with w >= 0
uint32_t quat_morton(const double* q) {
    const double s = q[0] < 0 ? -1.0 : 1.0;
    uint32_t m = 0, c[3];
    for (int a = 0; a < 3; a++) {
        const double v = std::min(1.0, std::max(-1.0, s * q[a + 1]));
        c[a] = (uint32_t)std::min(1023.0, std::floor((v + 1.0) * 512.0));
    }
    for (int b = 9; b >= 0; b--)
        for (int a = 0; a < 3; a++) m = (m << 1) | ((c[a] >> b) & 1u);
    return m;
}

std::vector<int> pose_order(const double* quat, const int* off, const int* cls, int P,
                            const int* cnt) {
    std::vector<std::pair<uint64_t, int>> key(P);
    for (int p = 0; p < P; p++)
        key[p] = {cnt && cnt[p] == 0   // no rotations: last, and its offset is not read
                      ? ~0ull
                      : ((uint64_t)(cls ? cls[p] : 0) << 32) | quat_morton(quat + 4 * (size_t)off[p]),
                  p};
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int>& x, const std::pair<uint64_t, int>& y) {
                         return x.first < y.first;
                     });
    std::vector<int> order(P);
    for (int i = 0; i < P; i++) order[i] = key[i].second;
    return order;
}

// Launch shape of the local scorer for this batch: the launch form (direct or
// persistent) and, in the 16-bit forms, the gather depth (4 or 8 samples in
// flight); each not requested is tuned: the first batch of each (precision,
// translation tile, power-of-two bucket of the item count) runs every
// candidate shape once, timed with events, and the fastest is kept.
int choose_local_mode(thx_ctx* c, int precision, int TPAD, int nblk, const LocalBatch& b) {
    std::vector<int> modes, nss;
    if (c->locModeReq == kLocalDirect || c->locModeReq == kLocalQueue) modes = {c->locModeReq};
    else modes = {kLocalDirect, kLocalQueue};
    if (precision == THX_FP32) nss = {4};
    else if (c->locNsReq == 4 || c->locNsReq == 8) nss = {c->locNsReq};
    else nss = {4, 8};
    std::vector<int> cand;
    for (int ns : nss)
        for (int m : modes) cand.push_back(local_shape(m, ns));
    if (cand.size() == 1) return cand[0];
    int lg = 0;
    while ((1 << lg) < nblk) lg++;
    const long long key = (((long long)precision * 1000 + TPAD) * 100 + lg) * 1000 +
                          c->locModeReq * 100 + c->locNsReq;
    for (auto& kv : c->locTuned)
        if (kv.first == key) return kv.second;
    const int n = (int)cand.size();
    std::vector<hipEvent_t> e(n + 1, nullptr);
    int s = THX_OK;
    for (int i = 0; i <= n && s == THX_OK; i++)
        if (hipEventCreate(&e[i]) != hipSuccess) s = THX_ERR_HIP;
    if (s == THX_OK && hipEventRecord(e[0], c->stream) != hipSuccess) s = THX_ERR_HIP;
    for (int i = 0; i < n && s == THX_OK; i++) {
        s = launch_score_local(c, precision, TPAD, nblk, b, cand[i]);
        if (s == THX_OK && hipEventRecord(e[i + 1], c->stream) != hipSuccess) s = THX_ERR_HIP;
    }
    if (s == THX_OK && hipEventSynchronize(e[n]) != hipSuccess) s = THX_ERR_HIP;
    int best = cand[0];
    float tBest = 0.f;
    for (int i = 0; i < n && s == THX_OK; i++) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e[i], e[i + 1]);
        if (i == 0 || t < tBest) { tBest = t; best = cand[i]; }
    }
    for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    if (s != THX_OK) {
        set_error("local scorer autotuning failed");
        return -1;
    }
    c->locTuned.push_back({key, best});
    return best;
}

static std::mutex g_default_mu;
static thx_ctx* g_default = nullptr;

static int default_ctx(thx_ctx** out) {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) THX_TRY(thx_ctx_create(0, &g_default));
    *out = g_default;
    return THX_OK;
}

// Bricked BF16 copies of the class volumes for the 16-bit local scorers:
// rebuilt on their next use after the volume changes.
void invalidate_vol_lo(thx_ctx* c, int cls) {
    if ((int)c->volLoOk.size() <= cls) c->volLoOk.resize(cls + 1, 0);
    c->volLoOk[cls] = 0;
}

int ensure_vol_lo(thx_ctx* c, const std::vector<char>& used) {
    const int nCls = (int)c->vol.size();
    if ((int)c->volLo.size() < nCls) c->volLo.resize(nCls);
    if ((int)c->volLoOk.size() < nCls) c->volLoOk.resize(nCls, 0);
    for (int k = 0; k < nCls; k++) {
        if (!used[k] || c->volLoOk[k]) continue;
        const int V = c->vdim[k];
        THX_TRY(c->volLo[k].ensure(vb_bytes(V)));
        THX_TRY(launch_vol_bricked(c->stream, c->vol[k].as<const float2>(), V,
                                   c->volLo[k].as<uint32_t>()));
        c->volLoOk[k] = 1;
    }
    return THX_OK;
}

// Particle input arrays for P particles (and the derived operands' rows).
int alloc_particles(thx_ctx* c, int P) {
    c->partFromImages = false;   // set again by the images path once it has packed them
    c->ctfReady = c->sigReady = c->ctfAttrOk = false;
    c->P = P;
    c->Ppad = round_up(P, 128);
    THX_TRY(c->dat.ensure(sizeof(float2) * (size_t)P * c->K));
    THX_TRY(c->ctf.ensure(sizeof(float) * (size_t)P * c->K));
    THX_TRY(c->sig.ensure(sizeof(float) * (size_t)P * c->K));
    THX_TRY(c->Aterm.ensure(sizeof(double) * c->Ppad));
    THX_TRY(c->Y.ensure(sizeof(float2) * (size_t)c->Ppad * c->Kp));
    THX_TRY(c->Wt.ensure(sizeof(float) * (size_t)c->Ppad * c->Kp));
    return THX_OK;
}

}  // namespace thx

using namespace thx;

extern "C" {

const char* thx_last_error(void) { return g_err.c_str(); }

int thx_device_count(int* n) {
    THX_CHECK(n, THX_ERR_ARG, "thx_device_count: null output");
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess) {
        *n = 0;
        set_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
        return THX_ERR_HIP;
    }
    *n = cnt;
    return THX_OK;
}

int thx_ctx_create(int device, thx_ctx** out) {
    clear_error();
    THX_CHECK(out, THX_ERR_ARG, "thx_ctx_create: null output");
    int n = 0;
    THX_TRY(thx_device_count(&n));
    THX_CHECK(device >= 0 && device < n, THX_ERR_ARG, "thx_ctx_create: no such device");
    thx_ctx* c = new thx_ctx();
    c->device = device;
    if (set_device(c) != THX_OK) { delete c; return THX_ERR_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        delete c;
        return THX_ERR_HIP;
    }
    for (int i = 0; i < 16; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            set_error("hipEventCreate failed");
            return THX_ERR_HIP;
        }
    *out = c;
    return THX_OK;
}

int thx_ctx_destroy(thx_ctx* c) {
    if (!c) return THX_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->locGraph) (void)hipGraphExecDestroy(c->locGraph);
    for (auto& g : c->pfGraphs) (void)hipGraphExecDestroy(g.second);
    c->pfGraphs.clear();
    c->pfCountsH.release();
    if (c->lpReady) (void)hipEventDestroy(c->lpReady);
    if (c->lpDone) (void)hipEventDestroy(c->lpDone);
    if (c->lpRead) (void)hipEventDestroy(c->lpRead);
    projectee_release(c);
    images_release(c);
    c->lpStage.release();
    DevBuf* bufs[] = {&c->iCol, &c->iRow, &c->iSig, &c->dat, &c->ctf, &c->sig, &c->Aterm,
                      &c->Y, &c->Wt, &c->quat, &c->trans, &c->pR, &c->pT, &c->pRf, &c->pTf,
                      &c->S, &c->rotP, &c->rotLo, &c->normP, &c->Cterm, &c->normPart, &c->shellScale,
                      &c->shellMax, &c->partMax, &c->partWC, &c->partArg, &c->partWR,
                      &c->partWT};
    for (DevBuf* b : bufs) b->release();
    DevBuf* lbufs[] = {&c->locQuat, &c->locTrans, &c->locPRf, &c->locPTf, &c->locROff,
                       &c->locNR, &c->locTOff, &c->locNT, &c->locBlocks, &c->locBStart,
                       &c->locPcls, &c->locVols, &c->locBMax, &c->locBWC, &c->locBArg,
                       &c->locBWT, &c->locWRrel, &c->locWC, &c->locWR, &c->locWT, &c->locBase,
                       &c->locBestR, &c->locBestT, &c->img, &c->lpMix, &c->lpRidx,
                       &c->lpRlogq, &c->lpTidx, &c->lpESS, &c->lpSpread, &c->lpTail,
                       &c->lpPmap, &c->lpOrder, &c->lpOrdKeyA, &c->lpOrdKeyB, &c->lpOrdVal,
                       &c->lpOrdTmp, &c->lpDat, &c->lpCtf, &c->lpSig, &c->lpQf, &c->S16,
                       &c->lpStageD,
                       &c->lpRes.wC,
                       &c->lpRes.wR, &c->lpRes.wT, &c->lpRes.base, &c->lpRes.bestR,
                       &c->lpRes.bestT, &c->lpRes.bestScore, &c->pfQ, &c->pfQ2, &c->pfTr,
                       &c->pfTr2, &c->pfWR, &c->pfWT, &c->pfWRf, &c->pfWTf, &c->pfIdx,
                       &c->pfWNew, &c->pfK, &c->pfS, &c->pfMean, &c->pfTopR, &c->pfTopT,
                       &c->pfPeak, &c->pfTransS, &c->pfBest, &c->pfActive, &c->pfNPhase,
                       &c->pfOffR, &c->pfOffT, &c->pfNR, &c->pfNT, &c->pfNRa, &c->pfNTa,
                       &c->pfOrder, &c->pfCounts,
                       &c->locCounter};
    for (DevBuf* b : lbufs) b->release();
    for (auto& v : c->vol) v.release();
    for (auto& v : c->volLo) v.release();
    c->locVolsLo.release();
    for (auto& r : c->res) {
        r.wC.release(); r.wR.release(); r.wT.release(); r.base.release();
        r.bestR.release(); r.bestT.release(); r.bestScore.release();
    }
    for (int i = 0; i < 16; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return THX_OK;
}

int thx_ctx_set_pixels(thx_ctx* c, const int* iCol, const int* iRow, const int* iSig,
                       int K, int idim, int pf) {
    clear_error();
    THX_CHECK(c && iCol && iRow, THX_ERR_ARG, "set_pixels: null argument");
    THX_CHECK(K > 0 && idim > 0 && pf > 0, THX_ERR_ARG, "set_pixels: bad size");
    THX_TRY(set_device(c));
    const int Kp = round_up(K, kKAlign);
    std::vector<int> col(Kp, 0), row(Kp, 0), sg(Kp, 0);
    int maxShell = 0;
    double maxR = 0;
    int maxCol = 0, maxAbsRow = 0;
    for (int k = 0; k < K; k++) {
        col[k] = iCol[k];
        row[k] = iRow[k];
        THX_CHECK(iCol[k] >= 0, THX_ERR_ARG, "set_pixels: iCol must be >= 0 (half plane)");
        const double rr = std::hypot((double)iCol[k], (double)iRow[k]);
        maxR = std::max(maxR, rr);
        sg[k] = iSig ? iSig[k] : (int)std::rint(rr);
        THX_CHECK(sg[k] >= 0, THX_ERR_ARG, "set_pixels: negative shell index");
        maxShell = std::max(maxShell, sg[k]);
        maxCol = std::max(maxCol, iCol[k]);
        maxAbsRow = std::max(maxAbsRow, std::abs(iRow[k]));
    }
    THX_CHECK(maxR * pf + 2 <= (double)idim * pf / 2, THX_ERR_RANGE,
              "set_pixels: pixel radius too close to Nyquist of the padded volume");
    THX_TRY(c->iCol.ensure(sizeof(int) * Kp));
    THX_TRY(c->iRow.ensure(sizeof(int) * Kp));
    THX_TRY(c->iSig.ensure(sizeof(int) * Kp));
    THX_HIP(hipMemcpy(c->iCol.p, col.data(), sizeof(int) * Kp, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->iRow.p, row.data(), sizeof(int) * Kp, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->iSig.p, sg.data(), sizeof(int) * Kp, hipMemcpyHostToDevice));
    c->K = K;
    c->Kp = Kp;
    c->idim = idim;
    c->pf = pf;
    c->maxShell = maxShell;
    c->pixMaxCol = maxCol;
    c->pixMaxAbsRow = maxAbsRow;
    c->P = 0;   // particle arrays depend on K
    c->R = c->T = 0;
    c->rotLoPrec = -1;   // tables derived from the projections are stale
    c->shellMaxPOk = false;
    return THX_OK;
}

int thx_ctx_set_volume(thx_ctx* c, int cls, const float* vol, int vdim) {
    clear_error();
    THX_CHECK(c && vol && cls >= 0 && vdim > 0 && vdim % 2 == 0, THX_ERR_ARG,
              "set_volume: bad argument");
    THX_TRY(set_device(c));
    if ((int)c->vol.size() <= cls) {
        c->vol.resize(cls + 1);
        c->vdim.resize(cls + 1, 0);
    }
    const size_t n = (size_t)(vdim / 2 + 1) * vdim * vdim;
    THX_TRY(c->vol[cls].ensure(sizeof(float2) * n));
    THX_HIP(hipMemcpy(c->vol[cls].p, vol, sizeof(float2) * n, hipMemcpyHostToDevice));
    c->vdim[cls] = vdim;
    invalidate_vol_lo(c, cls);
    return THX_OK;
}

int thx_ctx_get_volume(thx_ctx* c, int cls, float* vol) {
    clear_error();
    THX_CHECK(c && vol && cls >= 0 && cls < (int)c->vol.size() && c->vdim[cls] > 0,
              THX_ERR_STATE, "get_volume: no such class");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int V = c->vdim[cls];
    THX_HIP(hipMemcpy(vol, c->vol[cls].p, sizeof(float2) * (size_t)(V / 2 + 1) * V * V,
                      hipMemcpyDeviceToHost));
    return THX_OK;
}


int thx_ctx_set_particles(thx_ctx* c, const float* datP, const float* ctfP,
                          const float* sigRcpP, int P) {
    clear_error();
    THX_CHECK(c && datP && P > 0, THX_ERR_ARG, "set_particles: bad argument");
    THX_CHECK(c->K > 0, THX_ERR_STATE, "set_particles: pixel set not defined");
    THX_TRY(set_device(c));
    THX_TRY(alloc_particles(c, P));
    const size_t n = (size_t)P * c->K;
    THX_HIP(hipMemcpy(c->dat.p, datP, sizeof(float2) * n, hipMemcpyHostToDevice));
    if (ctfP) THX_HIP(hipMemcpy(c->ctf.p, ctfP, sizeof(float) * n, hipMemcpyHostToDevice));
    if (sigRcpP) THX_HIP(hipMemcpy(c->sig.p, sigRcpP, sizeof(float) * n, hipMemcpyHostToDevice));
    c->ctfReady = ctfP != nullptr;
    c->sigReady = sigRcpP != nullptr;
    return THX_OK;
}

int thx_ctx_set_ctf(thx_ctx* c, const float* ctfAttr, int P, float pixelSize) {
    clear_error();
    THX_CHECK(c && ctfAttr && P > 0 && pixelSize > 0.f, THX_ERR_ARG, "set_ctf: bad argument");
    THX_CHECK(c->K > 0 && c->P > 0, THX_ERR_STATE, "set_ctf: pixels and particles must be set");
    THX_CHECK(P == c->P, THX_ERR_ARG, "set_ctf: P differs from the context's particle count");
    for (int p = 0; p < P; p++)
        THX_CHECK(ctfAttr[7 * (size_t)p] > 0.f && ctfAttr[7 * (size_t)p + 5] >= 0.f &&
                      ctfAttr[7 * (size_t)p + 5] <= 1.f,
                  THX_ERR_ARG, "set_ctf: voltage must be > 0 and amplitude contrast in [0, 1]");
    THX_TRY(set_device(c));
    // the attributes stay on the device: the sigma refresh evaluates the CTF
    // over the whole spectrum from them
    THX_TRY(upload(c->ctfAttr, ctfAttr, 7 * (size_t)P, c->stream));
    int s = launch_ctf_attr(c, c->ctfAttr.as<const float>(), P, pixelSize);
    if (s == THX_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("set_ctf: kernel failed");
        s = THX_ERR_HIP;
    }
    if (s == THX_OK) c->ctfReady = c->ctfAttrOk = true;
    return s;
}

int thx_ctx_get_particles(thx_ctx* c, float* datP, float* ctfP, float* sigRcpP) {
    clear_error();
    THX_CHECK(c && c->P > 0, THX_ERR_STATE, "get_particles: no particles");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->P * c->K;
    if (datP) THX_HIP(hipMemcpy(datP, c->dat.p, sizeof(float2) * n, hipMemcpyDeviceToHost));
    if (ctfP) THX_HIP(hipMemcpy(ctfP, c->ctf.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (sigRcpP) THX_HIP(hipMemcpy(sigRcpP, c->sig.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return THX_OK;
}

static int set_grid_shape(thx_ctx* c, int R, int T) {
    c->R = R;
    c->T = T;
    c->Rpad = round_up(R, 256);
    c->TpadS = round_up(T, 128);
    return ensure_grid_buffers(c);
}

static int upload_priors(thx_ctx* c, const double* pR, const double* pT) {
    std::vector<float> fr(c->Rpad, 0.f), ft(c->TpadS, 0.f);
    for (int r = 0; r < c->R; r++) fr[r] = (float)pR[r];
    for (int t = 0; t < c->T; t++) ft[t] = (float)pT[t];
    THX_TRY(c->pRf.ensure(sizeof(float) * c->Rpad));
    THX_TRY(c->pTf.ensure(sizeof(float) * c->TpadS));
    THX_HIP(hipMemcpy(c->pRf.p, fr.data(), sizeof(float) * c->Rpad, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pTf.p, ft.data(), sizeof(float) * c->TpadS, hipMemcpyHostToDevice));
    // FP64 copies: the resampling prior of the particle filter
    THX_TRY(c->pR.ensure(sizeof(double) * c->R));
    THX_TRY(c->pT.ensure(sizeof(double) * c->T));
    THX_HIP(hipMemcpy(c->pR.p, pR, sizeof(double) * c->R, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->pT.p, pT, sizeof(double) * c->T, hipMemcpyHostToDevice));
    return THX_OK;
}

int thx_ctx_set_grid(thx_ctx* c, const double* quat, int R, const double* trans, int T,
                     const double* pR, const double* pT) {
    clear_error();
    THX_CHECK(c && quat && trans && pR && pT && R > 0 && T > 0, THX_ERR_ARG,
              "set_grid: bad argument");
    THX_CHECK(c->K > 0, THX_ERR_STATE, "set_grid: pixel set not defined");
    THX_TRY(set_device(c));
    THX_TRY(set_grid_shape(c, R, T));
    THX_TRY(c->quat.ensure(sizeof(double) * 4 * R));
    THX_TRY(c->trans.ensure(sizeof(double) * 2 * T));
    THX_HIP(hipMemcpy(c->quat.p, quat, sizeof(double) * 4 * R, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->trans.p, trans, sizeof(double) * 2 * T, hipMemcpyHostToDevice));
    THX_TRY(upload_priors(c, pR, pT));
    THX_TRY(launch_translate_table(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    return THX_OK;
}

int thx_ctx_score_global(thx_ctx* c, int cls, int precision) {
    clear_error();
    THX_CHECK(c, THX_ERR_ARG, "score_global: null context");
    THX_CHECK(precision >= THX_FP32 && precision <= THX_FP8, THX_ERR_ARG,
              "score_global: unknown precision");
    THX_CHECK(c->K > 0 && c->P > 0 && c->R > 0 && c->T > 0, THX_ERR_STATE,
              "score_global: pixels, particles and grid must be set");
    THX_CHECK(cls >= 0 && cls < (int)c->vol.size() && c->vdim[cls] > 0, THX_ERR_STATE,
              "score_global: volume of this class not set");
    THX_CHECK(c->vdim[cls] == c->idim * c->pf, THX_ERR_ARG,
              "score_global: volume size differs from idim * pf");
    THX_TRY(set_device(c));
    c->nev = 0;
    c->evNames.clear();
    THX_TRY(record(c, "start"));
    THX_TRY(launch_project(c, cls, c->quat.as<const double>(), c->R, c->Rpad,
                           c->rotP.as<float2>(), c->normP.as<float>()));
    THX_TRY(record(c, "project"));
    return run_scan(c, cls, precision);
}

int thx_ctx_sync(thx_ctx* c) {
    THX_CHECK(c, THX_ERR_ARG, "sync: null context");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    return THX_OK;
}

int thx_ctx_last_timing(thx_ctx* c, char* names, int cap, double* ms, int maxn, int* n) {
    clear_error();
    THX_CHECK(c && n, THX_ERR_ARG, "last_timing: null argument");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    std::string all;
    int cnt = 0;
    for (int i = 1; i < c->nev && cnt < maxn; i++) {
        float t = 0.f;
        THX_HIP(hipEventElapsedTime(&t, c->ev[i - 1], c->ev[i]));
        if (ms) ms[cnt] = t;
        if (!all.empty()) all += ";";
        all += c->evNames[i];
        cnt++;
    }
    if (names && cap > 0) {
        std::strncpy(names, all.c_str(), cap - 1);
        names[cap - 1] = 0;
    }
    *n = cnt;
    return THX_OK;
}

int thx_ctx_get_global(thx_ctx* c, int nK, double* wC, double* wR, double* wT, double* baseL) {
    clear_error();
    THX_CHECK(c && nK > 0 && wC && wR && wT && baseL, THX_ERR_ARG, "get_global: bad argument");
    THX_CHECK((int)c->res.size() >= nK, THX_ERR_STATE, "get_global: classes not scored");
    for (int k = 0; k < nK; k++)
        THX_CHECK(c->res[k].valid, THX_ERR_STATE, "get_global: a class was not scored");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int P = c->P, R = c->R, T = c->T;
    std::vector<double> cw(P), cb(P), cr((size_t)P * R), ct((size_t)P * T);
    for (int l = 0; l < P; l++) baseL[l] = -INFINITY;
    std::vector<std::vector<double>> bases(nK);
    for (int k = 0; k < nK; k++) {
        THX_HIP(hipMemcpy(cb.data(), c->res[k].base.p, sizeof(double) * P, hipMemcpyDeviceToHost));
        bases[k] = cb;
        for (int l = 0; l < P; l++) baseL[l] = std::max(baseL[l], cb[l]);
    }
    for (int k = 0; k < nK; k++) {
        const ClassResult& r = c->res[k];
        THX_HIP(hipMemcpy(cw.data(), r.wC.p, sizeof(double) * P, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(cr.data(), r.wR.p, sizeof(double) * P * R, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(ct.data(), r.wT.p, sizeof(double) * P * T, hipMemcpyDeviceToHost));
        for (int l = 0; l < P; l++) {
            const double g = std::exp(bases[k][l] - baseL[l]);
            wC[(size_t)l * nK + k] = g * cw[l];
            for (int i = 0; i < R; i++) wR[((size_t)l * nK + k) * R + i] = g * cr[(size_t)l * R + i];
            for (int i = 0; i < T; i++) wT[((size_t)l * nK + k) * T + i] = g * ct[(size_t)l * T + i];
        }
    }
    return THX_OK;
}

int thx_ctx_get_best(thx_ctx* c, int cls, int* bestR, int* bestT, double* bestScore) {
    clear_error();
    THX_CHECK(c && cls >= 0 && cls < (int)c->res.size() && c->res[cls].valid, THX_ERR_STATE,
              "get_best: class not scored");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    const ClassResult& r = c->res[cls];
    if (bestR) THX_HIP(hipMemcpy(bestR, r.bestR.p, sizeof(int) * c->P, hipMemcpyDeviceToHost));
    if (bestT) THX_HIP(hipMemcpy(bestT, r.bestT.p, sizeof(int) * c->P, hipMemcpyDeviceToHost));
    if (bestScore)
        THX_HIP(hipMemcpy(bestScore, r.bestScore.p, sizeof(double) * c->P, hipMemcpyDeviceToHost));
    return THX_OK;
}

int thx_ctx_score_one(thx_ctx* c, int cls, int p, float* out) {
    clear_error();
    THX_CHECK(c && out, THX_ERR_ARG, "score_one: null argument");
    THX_CHECK(c->K > 0 && c->P > 0 && c->R > 0 && c->T > 0, THX_ERR_STATE,
              "score_one: pixels, particles and grid must be set");
    THX_CHECK(p >= 0 && p < c->P, THX_ERR_ARG, "score_one: particle out of range");
    THX_CHECK(cls >= 0 && cls < (int)c->vol.size() && c->vdim[cls] > 0, THX_ERR_STATE,
              "score_one: volume not set");
    THX_TRY(set_device(c));
    THX_TRY(launch_project(c, cls, c->quat.as<const double>(), c->R, c->Rpad,
                           c->rotP.as<float2>(), c->normP.as<float>()));
    DevBuf tmp;
    THX_TRY(tmp.ensure(sizeof(float) * (size_t)c->R * c->T));
    int s = launch_score_one(c, cls, p, tmp.as<float>());
    if (s == THX_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess)
            e = hipMemcpy(out, tmp.p, sizeof(float) * (size_t)c->R * c->T, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error(std::string("score_one: ") + hipGetErrorString(e));
            s = THX_ERR_HIP;
        }
    }
    tmp.release();
    return s;
}

int thx_ctx_score_local(thx_ctx* c, const int* cls, const double* quat, const int* nR,
                        const double* trans, const int* nT, const double* pR, const double* pT,
                        int precision) {
    clear_error();
    THX_CHECK(c && quat && nR && trans && nT && pR && pT, THX_ERR_ARG,
              "score_local: null argument");
    THX_CHECK(precision == THX_FP32 || precision == THX_BF16 || precision == THX_FP16,
              THX_ERR_ARG, "score_local: precision must be FP32, BF16 or FP16");
    THX_CHECK(c->K > 0 && c->P > 0, THX_ERR_STATE, "score_local: pixels and particles must be set");
    THX_CHECK(!c->vol.empty(), THX_ERR_STATE, "score_local: no volume set");
    const int P = c->P;
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
                  "score_local: a class volume in use does not match idim * pf");
    std::vector<int> rOff(P), tOff(P), bStart(P + 1);
    std::vector<int2> blocks;
    size_t Rtot = 0, Ttot = 0;
    int maxT = 0;
    int nScored = 0;
    for (int p = 0; p < P; p++) {
        // nR[p] == 0 skips particle p (wC 0, baseL -inf, no best pose)
        THX_CHECK(nR[p] >= 0 && nR[p] <= 64 * kLocalRotBlock, THX_ERR_RANGE,
                  "score_local: rotations per particle must be in [0, 8192]");
        THX_CHECK((nR[p] == 0 && nT[p] >= 0 && nT[p] <= 128) || (nT[p] >= 1 && nT[p] <= 128),
                  THX_ERR_RANGE, "score_local: translations per particle must be in [1, 128]");
        nScored += nR[p] > 0;
        if (cls) THX_CHECK(cls[p] >= 0 && cls[p] < nCls, THX_ERR_ARG, "score_local: bad class");
        rOff[p] = (int)Rtot;
        tOff[p] = (int)Ttot;
        Rtot += nR[p];
        Ttot += nT[p];
        maxT = std::max(maxT, nT[p]);
    }
    THX_CHECK(nScored > 0, THX_ERR_ARG, "score_local: no particle has rotations to score");
    // work items particle by particle in pose order (class, then first
    // rotation), grouped by translation tile: particles with at most 32, 64 and
    // 128 translations each go to their own scorer launch with that tile, so a
    // ragged batch pays for its padding per particle, not for its largest T_p
    std::fill(bStart.begin(), bStart.end(), 0);
    const std::vector<int> order = pose_order(quat, rOff.data(), cls, P, nR);
    int tileStart[kLocalTiles + 1];
    for (int ti = 0; ti < kLocalTiles; ti++) {
        tileStart[ti] = (int)blocks.size();
        for (int p : order) {
            if (nR[p] == 0 || local_tile_index(nT[p]) != ti) continue;
            bStart[p] = (int)blocks.size();
            for (int r0 = 0; r0 < nR[p]; r0 += kLocalRotBlock) blocks.push_back(make_int2(p, r0));
        }
    }
    tileStart[kLocalTiles] = (int)blocks.size();
    bStart[P] = (int)blocks.size();
    THX_CHECK(Rtot < (1u << 31) && Ttot < (1u << 31), THX_ERR_RANGE, "score_local: batch too large");
    const int TPAD = kLocalTile[local_tile_index(maxT)];
    const int nblk = (int)blocks.size();
    // each particle's rotations in pose order, so the 128 rotations of a work
    // item are a compact cluster and their gathers share cache lines; wR and
    // the best rotation are mapped back by thx_ctx_get_local
    c->locPerm.resize(Rtot);
    c->locROffH = rOff;
    std::vector<double> qP(4 * Rtot);
    std::vector<float> fR(Rtot), fT(Ttot);
    {
        // particles split over host threads (large batches take seconds serially)
        auto sort_range = [&](int pa, int pb) {
            std::vector<std::pair<uint32_t, int>> key;
            for (int p = pa; p < pb; p++) {
                key.resize(nR[p]);
                for (int r = 0; r < nR[p]; r++)
                    key[r] = {quat_morton(quat + 4 * ((size_t)rOff[p] + r)), rOff[p] + r};
                std::sort(key.begin(), key.end());
                for (int r = 0; r < nR[p]; r++) {
                    const size_t dst = (size_t)rOff[p] + r, src = (size_t)key[r].second;
                    c->locPerm[dst] = (int)src;
                    for (int e = 0; e < 4; e++) qP[4 * dst + e] = quat[4 * src + e];
                    fR[dst] = (float)pR[src];
                }
            }
        };
        const int nth = Rtot < (1u << 20) ? 1
                        : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        int started = 1;   // ranges handed to threads: 1 .. started-1
        try {
            for (; started < nth; started++)
                pool.emplace_back(sort_range, (int)((long long)P * started / nth),
                                  (int)((long long)P * (started + 1) / nth));
        } catch (...) {
            // no more threads: the remaining ranges run here
        }
        sort_range(0, (int)((long long)P / nth));
        for (int i = started; i < nth; i++)
            sort_range((int)((long long)P * i / nth), (int)((long long)P * (i + 1) / nth));
        for (auto& t : pool) t.join();
    }
    for (size_t i = 0; i < Ttot; i++) fT[i] = (float)pT[i];
    std::vector<const float2*> vp(nCls);
    for (int k = 0; k < nCls; k++) vp[k] = c->vol[k].as<const float2>();
    THX_TRY(set_device(c));
    hipStream_t st = c->stream;
    THX_TRY(upload(c->locQuat, qP.data(), 4 * Rtot, st));
    THX_TRY(upload(c->locTrans, trans, 2 * Ttot, st));
    THX_TRY(upload(c->locPRf, fR.data(), Rtot, st));
    THX_TRY(upload(c->locPTf, fT.data(), Ttot, st));
    THX_TRY(upload(c->locROff, rOff.data(), P, st));
    THX_TRY(upload(c->locNR, nR, P, st));
    THX_TRY(upload(c->locTOff, tOff.data(), P, st));
    THX_TRY(upload(c->locNT, nT, P, st));
    THX_TRY(upload(c->locBlocks, blocks.data(), blocks.size(), st));
    THX_TRY(upload(c->locBStart, bStart.data(), P + 1, st));
    THX_TRY(upload(c->locVols, vp.data(), vp.size(), st));
    if (precision != THX_FP32) {
        THX_TRY(ensure_vol_lo(c, used));
        std::vector<const uint32_t*> vl(nCls, nullptr);
        for (int k = 0; k < nCls; k++)
            if (used[k]) vl[k] = c->volLo[k].as<const uint32_t>();
        THX_TRY(upload(c->locVolsLo, vl.data(), vl.size(), st));
    }
    if (cls) THX_TRY(upload(c->locPcls, cls, P, st));
    THX_TRY(c->locBMax.ensure(sizeof(float) * nblk));
    THX_TRY(c->locBWC.ensure(sizeof(float) * nblk));
    THX_TRY(c->locBArg.ensure(sizeof(int) * nblk));
    THX_TRY(c->locBWT.ensure(sizeof(float) * (size_t)nblk * kLocalWtStride));
    THX_TRY(c->locWRrel.ensure(sizeof(float) * Rtot));
    THX_TRY(c->locWC.ensure(sizeof(double) * P));
    THX_TRY(c->locWR.ensure(sizeof(double) * Rtot));
    THX_TRY(c->locWT.ensure(sizeof(double) * Ttot));
    THX_TRY(c->locBase.ensure(sizeof(double) * P));
    THX_TRY(c->locBestR.ensure(sizeof(int) * P));
    THX_TRY(c->locBestT.ensure(sizeof(int) * P));
    c->locP = P;
    c->locTPAD = TPAD;
    c->locNblk = nblk;
    c->locRtot = Rtot;
    c->locTtot = Ttot;
    LocalBatch b;
    b.vols = c->locVols.as<const float2* const>();
    b.volsLo = precision != THX_FP32 ? c->locVolsLo.as<const uint32_t* const>() : nullptr;
    b.pcls = cls ? c->locPcls.as<const int>() : nullptr;
    b.quat = c->locQuat.as<const double>();
    b.pRf = c->locPRf.as<const float>();
    b.rOff = c->locROff.as<const int>();
    b.nR = c->locNR.as<const int>();
    b.trans = c->locTrans.as<const double>();
    b.pTf = c->locPTf.as<const float>();
    b.tOff = c->locTOff.as<const int>();
    b.nT = c->locNT.as<const int>();
    b.blocks = c->locBlocks.as<const int2>();
    b.bStart = c->locBStart.as<const int>();
    b.bMax = c->locBMax.as<float>();
    b.bWC = c->locBWC.as<float>();
    b.bArg = c->locBArg.as<int>();
    b.bWT = c->locBWT.as<float>();
    b.wRrel = c->locWRrel.as<float>();
    c->nev = 0;
    c->evNames.clear();
    // FP16: per-shell operand scales from the class volumes in use and the
    // particles (eager: no memset node may enter a captured graph)
    const int prepPrec = precision == THX_FP16 ? THX_FP16 : THX_FP32;
    if (precision == THX_FP16) THX_TRY(launch_local_shell_scale(c, used));
    THX_TRY(launch_prep_particles(c, prepPrec));
    // launch shape per tile class (the tuner's choice, or the requested one)
    int modes[kLocalTiles] = {};
    int most = 0;
    for (int ti = 0; ti < kLocalTiles; ti++) {
        const int n = tileStart[ti + 1] - tileStart[ti];
        if (n == 0) continue;
        const LocalBatch bt = local_batch_range(b, tileStart[ti]);
        modes[ti] = choose_local_mode(c, precision, kLocalTile[ti], n, bt);
        if (modes[ti] < 0) return THX_ERR_HIP;
        if (n > most) { most = n; c->locModeUsed = modes[ti]; }
    }
    // the scorer launches of all tile classes
    auto score_all = [&]() -> int {
        for (int ti = 0; ti < kLocalTiles; ti++) {
            const int n = tileStart[ti + 1] - tileStart[ti];
            if (n == 0) continue;
            THX_TRY(launch_score_local(c, precision, kLocalTile[ti], n,
                                       local_batch_range(b, tileStart[ti]), modes[ti]));
        }
        return THX_OK;
    };
    if (c->locGraphOn) {
        THX_TRY(c->locCounter.ensure(8 * sizeof(int)));   // no allocation inside the capture
        // one captured graph per batch shape and buffer set: operand prep,
        // scorer, merge; replayed while they stay the same
        auto ptrs = [](std::initializer_list<const void*> l) {
            std::vector<long long> v;
            for (const void* x : l) v.push_back((long long)(uintptr_t)x);
            return v;
        };
        std::vector<long long> key = ptrs(
            {c->Y.p, c->Wt.p, c->Aterm.p, c->dat.p, c->ctf.p, c->sig.p, c->iCol.p, c->iRow.p,
             c->iSig.p, c->locQuat.p, c->locTrans.p, c->locPRf.p, c->locPTf.p, c->locROff.p,
             c->locNR.p, c->locTOff.p, c->locNT.p, c->locBlocks.p, c->locBStart.p, c->locVols.p,
             precision != THX_FP32 ? c->locVolsLo.p : nullptr,
             cls ? c->locPcls.p : nullptr, c->locBMax.p, c->locBWC.p, c->locBArg.p, c->locBWT.p,
             c->locWRrel.p, c->locWC.p, c->locWR.p, c->locWT.p, c->locBase.p, c->locBestR.p,
             c->locBestT.p, c->locCounter.p});
        const long long shape[] = {P, nblk, precision, c->K, c->Kp, c->idim};
        key.insert(key.end(), shape, shape + 6);
        for (int ti = 0; ti < kLocalTiles; ti++) {
            key.push_back(tileStart[ti + 1] - tileStart[ti]);
            key.push_back(modes[ti]);
        }
        if (!c->locGraph || key != c->locGraphKey) {
            if (c->locGraph) {
                (void)hipGraphExecDestroy(c->locGraph);
                c->locGraph = nullptr;
            }
            THX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int cs = launch_prep_particles(c, prepPrec);
            if (cs == THX_OK) cs = score_all();
            if (cs == THX_OK)
                cs = launch_merge_local(c, P, kLocalWtStride, b, c->locWC.as<double>(),
                                        c->locWR.as<double>(), c->locWT.as<double>(),
                                        c->locBase.as<double>(), c->locBestR.as<int>(),
                                        c->locBestT.as<int>());
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(st, &g);
            if (cs != THX_OK) {
                if (g) (void)hipGraphDestroy(g);
                return cs;
            }
            THX_CHECK(ee == hipSuccess && g, THX_ERR_HIP, "score_local: graph capture failed");
            const hipError_t ie = hipGraphInstantiate(&c->locGraph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            THX_CHECK(ie == hipSuccess, THX_ERR_HIP, "score_local: graph instantiation failed");
            c->locGraphKey = key;
        }
        THX_TRY(record(c, "start"));
        THX_HIP(hipGraphLaunch(c->locGraph, st));
        THX_TRY(record(c, "local_graph"));
        c->locValid = true;
        return THX_OK;
    }
    THX_TRY(record(c, "start"));
    THX_TRY(launch_prep_particles(c, prepPrec));
    THX_TRY(record(c, "prep"));
    THX_TRY(score_all());
    THX_TRY(record(c, "score_local"));
    THX_TRY(launch_merge_local(c, P, kLocalWtStride, b, c->locWC.as<double>(), c->locWR.as<double>(),
                               c->locWT.as<double>(), c->locBase.as<double>(),
                               c->locBestR.as<int>(), c->locBestT.as<int>()));
    THX_TRY(record(c, "merge_local"));
    c->locValid = true;
    return THX_OK;
}

int thx_ctx_set_local_options(thx_ctx* c, int mode, int graph) {
    clear_error();
    THX_CHECK(c && mode >= 0 && mode <= 2, THX_ERR_ARG, "set_local_options: bad argument");
    c->locModeReq = mode;
    c->locGraphOn = graph != 0;
    return THX_OK;
}

int thx_ctx_set_local_gather(thx_ctx* c, int ns) {
    clear_error();
    THX_CHECK(c && (ns == 0 || ns == 4 || ns == 8), THX_ERR_ARG,
              "set_local_gather: samples in flight must be 0 (tuned), 4 or 8");
    c->locNsReq = ns;
    return THX_OK;
}

int thx_ctx_set_learned_fallback_selection(thx_ctx* c, int nSelR, int nSelT) {
    clear_error();
    THX_CHECK(c && nSelR >= 0 && nSelT >= 0 && (nSelR == 0) == (nSelT == 0), THX_ERR_ARG,
              "set_learned_fallback_selection: sizes must both be positive or both 0");
    THX_CHECK(nSelR <= 64 * 2 * kLocalRotBlock && nSelT <= 128, THX_ERR_RANGE,
              "set_learned_fallback_selection: at most 16384 rotations and 128 translations");
    c->lpWideR = nSelR;
    c->lpWideT = nSelT;
    return THX_OK;
}

int thx_ctx_set_select_order(thx_ctx* c, int on) {
    clear_error();
    THX_CHECK(c && (on == 0 || on == 1), THX_ERR_ARG, "set_select_order: on must be 0 or 1");
    c->selOrder = on != 0;
    return THX_OK;
}

int thx_ctx_local_mode(thx_ctx* c, int* mode) {
    clear_error();
    THX_CHECK(c && mode, THX_ERR_ARG, "local_mode: bad argument");
    *mode = c->locModeUsed;
    return THX_OK;
}

int thx_ctx_get_local(thx_ctx* c, double* wC, double* wR, double* wT, double* baseL,
                      int* bestR, int* bestT) {
    clear_error();
    THX_CHECK(c && c->locValid, THX_ERR_STATE, "get_local: no local scan done");
    THX_TRY(set_device(c));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int P = c->locP;
    if (wC) THX_HIP(hipMemcpy(wC, c->locWC.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    if (wR) {
        // the scorer's rotations were in pose order: back to the caller's
        std::vector<double> w(c->locRtot);
        THX_HIP(hipMemcpy(w.data(), c->locWR.p, sizeof(double) * c->locRtot, hipMemcpyDeviceToHost));
        const bool perm = c->locPerm.size() == c->locRtot;
        for (size_t i = 0; i < c->locRtot; i++) wR[perm ? c->locPerm[i] : i] = w[i];
    }
    if (wT) THX_HIP(hipMemcpy(wT, c->locWT.p, sizeof(double) * c->locTtot, hipMemcpyDeviceToHost));
    if (baseL) THX_HIP(hipMemcpy(baseL, c->locBase.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    if (bestR) {
        THX_HIP(hipMemcpy(bestR, c->locBestR.p, sizeof(int) * P, hipMemcpyDeviceToHost));
        if (c->locPerm.size() == c->locRtot && (int)c->locROffH.size() == P)
            for (int p = 0; p < P; p++)
                if (bestR[p] >= 0)
                    bestR[p] = c->locPerm[(size_t)c->locROffH[p] + bestR[p]] - c->locROffH[p];
    }
    if (bestT) THX_HIP(hipMemcpy(bestT, c->locBestT.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Learned-proposal expectation (grid-restricted)
// ---------------------------------------------------------------------------

// Scores nP particles (rows pmap or 0..nP-1) against per-particle selections
// (ridx / tidx, null = whole grid) with the gather scorer and merges in FP64
// into the loc* result buffers.  `order` (nP particle indices, or null) is the
// order in which the blocks take the particles; the results do not depend on it.
static int score_selected(thx_ctx* c, int nP, const int* ridx, const int* tidx, int nSelR,
                          int nSelT, const int* pmap, const int* order) {
    const int TPAD = nSelT <= 32 ? 32 : (nSelT <= 64 ? 64 : 128);
    // 256 rotations per block share one staged Z tile when there are enough
    const int RB = nSelR > kLocalRotBlock ? 2 * kLocalRotBlock : kLocalRotBlock;
    const int nbp = (nSelR + RB - 1) / RB;
    THX_CHECK(nbp <= 64, THX_ERR_RANGE, "learned: too many selected rotations");
    const size_t nblk = (size_t)nP * nbp;
    THX_CHECK(nblk < (1u << 31), THX_ERR_RANGE, "learned: grid too large");
    THX_TRY(c->locBMax.ensure(sizeof(float) * nblk));
    THX_TRY(c->locBWC.ensure(sizeof(float) * nblk));
    THX_TRY(c->locBArg.ensure(sizeof(int) * nblk));
    THX_TRY(c->locBWT.ensure(sizeof(float) * nblk * TPAD));
    THX_TRY(c->locWRrel.ensure(sizeof(float) * (size_t)nP * nSelR));
    THX_TRY(c->locWC.ensure(sizeof(double) * nP));
    THX_TRY(c->locWR.ensure(sizeof(double) * (size_t)nP * nSelR));
    THX_TRY(c->locWT.ensure(sizeof(double) * (size_t)nP * nSelT));
    THX_TRY(c->locBase.ensure(sizeof(double) * nP));
    THX_TRY(c->locBestR.ensure(sizeof(int) * nP));
    THX_TRY(c->locBestT.ensure(sizeof(int) * nP));
    // every particle has nSelR rotations and nSelT translations: the merge
    // takes the affine offsets p * nSelR, p * nSelT and blocks p * nbp
    LocalBatch b;
    b.affR = nSelR;
    b.affT = nSelT;
    b.affB = nbp;
    b.bMax = c->locBMax.as<float>();
    b.bWC = c->locBWC.as<float>();
    b.bArg = c->locBArg.as<int>();
    b.bWT = c->locBWT.as<float>();
    b.wRrel = c->locWRrel.as<float>();
    // the phase table as the scorer reads it (FP16), from the current S
    THX_TRY(launch_phase_half(c));
    THX_TRY(launch_score_sel(c, TPAD, RB, (int)nblk, ridx, tidx, nSelR, nSelT, 1.f / c->R,
                             1.f / c->T, pmap, order, b));
    THX_TRY(launch_merge_local(c, nP, TPAD, b, c->locWC.as<double>(), c->locWR.as<double>(),
                               c->locWT.as<double>(), c->locBase.as<double>(),
                               c->locBestR.as<int>(), c->locBestT.as<int>(), RB));
    return THX_OK;
}

// Order in which the selected-pose scorer takes the n particles whose
// mixtures are `mix` (proposal-mode order; null when ordering is off).
static int select_order(thx_ctx* c, const double* mix, int C, int n, const int** order) {
    *order = nullptr;
    if (!c->selOrder) return THX_OK;
    // particles in proposal-mode order: blocks resident together share rows
    size_t tmpBytes = 0;
    THX_TRY(pf_select_order(c->stream, nullptr, C, n, nullptr, nullptr, nullptr, nullptr,
                            nullptr, &tmpBytes));
    THX_TRY(c->lpOrder.ensure(sizeof(int) * n));
    THX_TRY(c->lpOrdKeyA.ensure(sizeof(uint32_t) * n));
    THX_TRY(c->lpOrdKeyB.ensure(sizeof(uint32_t) * n));
    THX_TRY(c->lpOrdVal.ensure(sizeof(int) * n));
    THX_TRY(c->lpOrdTmp.ensure(tmpBytes));
    THX_TRY(pf_select_order(c->stream, mix, C, n, c->lpOrdKeyA.as<uint32_t>(),
                            c->lpOrdKeyB.as<uint32_t>(), c->lpOrdVal.as<int>(),
                            c->lpOrder.as<int>(), c->lpOrdTmp.p, &tmpBytes));
    *order = c->lpOrder.as<int>();
    return THX_OK;
}

// Fallback by a wider selection: the n flagged particles (pinned indices
// idx) are rescored over the lpWideR x lpWideT most probable grid poses of
// their proposals with the selected-pose scorer; best pose (grid indices) and
// score land in locBestR / locBestT / locBase, rows in the order of idx.
static int rescore_wide(thx_ctx* c, int C, int n, const int* pinnedIdx) {
    hipStream_t st = c->stream;
    const int nR2 = std::min(c->lpWideR, c->R), nT2 = std::min(c->lpWideT, c->T);
    const int mixLen = 6 * C + 3;
    THX_TRY(c->lpPmap.ensure(sizeof(int) * n));
    THX_HIP(hipMemcpyAsync(c->lpPmap.p, pinnedIdx, sizeof(int) * n, hipMemcpyHostToDevice, st));
    THX_TRY(c->lpMixSub.ensure(sizeof(double) * (size_t)n * mixLen));
    THX_TRY(launch_gather_rows(st, c->lpMix.as<double>(), mixLen, c->lpPmap.as<int>(), n,
                               c->lpMixSub.as<double>()));
    THX_TRY(c->lpRidx.ensure(sizeof(int) * (size_t)n * nR2));
    THX_TRY(c->lpRlogq.ensure(sizeof(float) * (size_t)n * nR2));
    THX_TRY(c->lpTidx.ensure(sizeof(int) * (size_t)n * nT2));
    THX_TRY(c->lpSpread.ensure(sizeof(float) * n));
    THX_TRY(pf_select_grid(st, c->lpMixSub.as<double>(), C, n, c->quat.as<double>(), c->R,
                           c->trans.as<double>(), c->T, nR2, nT2, c->lpRidx.as<int>(),
                           c->lpRlogq.as<float>(), c->lpTidx.as<int>(), c->lpSpread.as<float>(),
                           c->lpQf.as<float4>()));
    const int* order = nullptr;
    THX_TRY(select_order(c, c->lpMixSub.as<double>(), C, n, &order));
    THX_TRY(record(c, "fallback_select"));
    THX_TRY(score_selected(c, n, c->lpRidx.as<int>(), c->lpTidx.as<int>(), nR2, nT2,
                           c->lpPmap.as<int>(), order));
    THX_TRY(c->lpESS.ensure(sizeof(double) * n));
    THX_TRY(c->lpTail.ensure(sizeof(float) * n));
    THX_TRY(launch_sel_stats(st, c->locWR.as<double>(), c->lpRlogq.as<float>(),
                             c->lpRidx.as<int>(), c->lpTidx.as<int>(), c->locBestR.as<int>(),
                             c->locBestT.as<int>(), nR2, nT2, n, c->lpESS.as<double>(),
                             c->lpTail.as<float>()));
    THX_TRY(record(c, "fallback_score"));
    return THX_OK;
}

// Whole-grid scan (run_scan) of the particles idx: their rows are gathered,
// and the context's particle inputs and this class's result slot are
// swapped for the gathered rows and lpRes for the duration of the scan.
}  // extern "C"

int thx::scan_subset(thx_ctx* c, int cls, int precision, const std::vector<int>& idx,
                     const int* pinnedIdx) {
    const int n = (int)idx.size();
    const size_t nk = (size_t)n * c->K;
    if (pinnedIdx) {   // a pinned copy of idx: the upload does not block the host
        THX_TRY(c->lpPmap.ensure(sizeof(int) * n));
        THX_HIP(hipMemcpyAsync(c->lpPmap.p, pinnedIdx, sizeof(int) * n, hipMemcpyHostToDevice,
                               c->stream));
    } else {
        THX_TRY(upload(c->lpPmap, idx.data(), n, c->stream));
    }
    THX_TRY(c->lpDat.ensure(sizeof(float2) * nk));
    THX_TRY(c->lpCtf.ensure(sizeof(float) * nk));
    THX_TRY(c->lpSig.ensure(sizeof(float) * nk));
    THX_TRY(launch_gather_particles(c, c->lpPmap.as<int>(), n, c->lpDat.as<float2>(),
                                    c->lpCtf.as<float>(), c->lpSig.as<float>()));
    if ((int)c->res.size() <= cls) c->res.resize(cls + 1);
    const int P0 = c->P, Ppad0 = c->Ppad;
    auto swap_in = [&]() {
        std::swap(c->dat, c->lpDat);
        std::swap(c->ctf, c->lpCtf);
        std::swap(c->sig, c->lpSig);
        std::swap(c->res[cls], c->lpRes);
    };
    swap_in();
    c->P = n;
    c->Ppad = round_up(n, 128);
    const int s = run_scan(c, cls, precision);
    swap_in();
    c->P = P0;
    c->Ppad = Ppad0;
    return s;
}

extern "C" {

int thx_ctx_expect_learned(thx_ctx* c, thx_cnn* net, int cls, int nSelR, int nSelT, double essMin,
                           double spreadMax, double tailMax, double scoreDrop,
                           int fallbackPrecision, int* nFallback) {
    clear_error();
    THX_CHECK(c && net && nSelR > 0 && nSelT > 0, THX_ERR_ARG, "expect_learned: bad argument");
    THX_CHECK(fallbackPrecision >= THX_FP32 && fallbackPrecision <= THX_FP8, THX_ERR_ARG,
              "expect_learned: unknown fallback precision");
    THX_CHECK(c->K > 0 && c->P > 0 && c->R > 0 && c->T > 0, THX_ERR_STATE,
              "expect_learned: pixels, particles and grid must be set");
    THX_CHECK(cls >= 0 && cls < (int)c->vol.size() && c->vdim[cls] == c->idim * c->pf,
              THX_ERR_STATE, "expect_learned: volume of this class not set");
    THX_CHECK(nSelR <= c->R && nSelT <= c->T && nSelT <= 128 && c->R <= 16384 && c->T <= 512,
              THX_ERR_RANGE, "expect_learned: selection sizes out of range");
    int C = 0;
    THX_TRY(thx_cnn_components(net, &C));
    THX_TRY(set_device(c));
    const int P = c->P;
    hipStream_t st = c->stream;
    c->nev = 0;
    c->evNames.clear();
    THX_TRY(c->lpMix.ensure(sizeof(double) * (size_t)P * (6 * C + 3)));
    THX_TRY(c->rotLo.ensure(sizeof(uint32_t) * (size_t)c->Rpad * c->Kp));
    if (!c->lpReady) THX_HIP(hipEventCreateWithFlags(&c->lpReady, hipEventDisableTiming));
    if (!c->lpDone) THX_HIP(hipEventCreateWithFlags(&c->lpDone, hipEventDisableTiming));
    THX_TRY(record(c, "start"));
    // the proposal network (images, layers, head) on its own stream, while the
    // context stream projects the grid and prepares the particle operands
    THX_HIP(hipEventRecord(c->lpReady, st));
    THX_TRY(cnn_propose_async(net, c, c->lpMix.as<double>(), c->lpReady, c->lpDone));
    // slices of the grid, once for all particles; BF16 table (written in the
    // same pass) + FP32 |P|^2
    THX_TRY(launch_project(c, cls, c->quat.as<const double>(), c->R, c->Rpad,
                           c->rotP.as<float2>(), c->normP.as<float>(), c->rotLo.as<uint32_t>()));
    THX_TRY(record(c, "project"));
    THX_TRY(launch_prep_particles(c, THX_FP32));
    THX_TRY(record(c, "prep"));
    // the rest of the network's time (what the projection did not cover)
    THX_HIP(hipStreamWaitEvent(st, c->lpDone, 0));
    THX_TRY(record(c, "propose"));
    THX_TRY(c->lpRidx.ensure(sizeof(int) * (size_t)P * nSelR));
    THX_TRY(c->lpRlogq.ensure(sizeof(float) * (size_t)P * nSelR));
    THX_TRY(c->lpTidx.ensure(sizeof(int) * (size_t)P * nSelT));
    THX_TRY(c->lpSpread.ensure(sizeof(float) * P));
    THX_TRY(c->lpQf.ensure(sizeof(float4) * c->R));
    THX_TRY(pf_select_grid(st, c->lpMix.as<double>(), C, P, c->quat.as<double>(), c->R,
                           c->trans.as<double>(), c->T, nSelR, nSelT, c->lpRidx.as<int>(),
                           c->lpRlogq.as<float>(), c->lpTidx.as<int>(), c->lpSpread.as<float>(),
                           c->lpQf.as<float4>()));
    const int* order = nullptr;
    THX_TRY(select_order(c, c->lpMix.as<double>(), C, P, &order));
    THX_TRY(record(c, "select"));
    THX_TRY(score_selected(c, P, c->lpRidx.as<int>(), c->lpTidx.as<int>(), nSelR, nSelT, nullptr,
                           order));
    THX_TRY(c->lpESS.ensure(sizeof(double) * P));
    THX_TRY(c->lpTail.ensure(sizeof(float) * P));
    THX_TRY(launch_sel_stats(st, c->locWR.as<double>(), c->lpRlogq.as<float>(),
                             c->lpRidx.as<int>(), c->lpTidx.as<int>(), c->locBestR.as<int>(),
                             c->locBestT.as<int>(), nSelR, nSelT, P, c->lpESS.as<double>(),
                             c->lpTail.as<float>()));
    THX_TRY(record(c, "score_sel"));
    // best pose in grid indices (mapped by the stats kernel)
    // the six per-particle results in one pinned batch: [score, ess] as
    // doubles, then [bestR, bestT] as ints and [spread, tail] as floats
    THX_TRY(c->lpStage.ensure((2 * sizeof(double) + 4 * sizeof(int)) * (size_t)P));
    double* hScore = c->lpStage.as<double>();
    double* hEss = hScore + P;
    int* hR = reinterpret_cast<int*>(hEss + P);
    int* hT = hR + P;
    float* hSpread = reinterpret_cast<float*>(hT + P);
    float* hTail = hSpread + P;
    // packed on the device in the staging layout: one copy instead of six
    const size_t stageBytes = (2 * sizeof(double) + 4 * sizeof(int)) * (size_t)P;
    THX_TRY(c->lpStageD.ensure(stageBytes));
    THX_TRY(launch_pack_learned(st, c->locBase.as<const double>(), c->lpESS.as<const double>(),
                                c->locBestR.as<const int>(), c->locBestT.as<const int>(),
                                c->lpSpread.as<const float>(), c->lpTail.as<const float>(), P,
                                c->lpStageD.p));
    THX_HIP(hipMemcpyAsync(hScore, c->lpStageD.p, stageBytes, hipMemcpyDeviceToHost, st));
    if (!c->lpRead) THX_HIP(hipEventCreateWithFlags(&c->lpRead, hipEventDisableTiming));
    THX_HIP(hipEventRecord(c->lpRead, st));
    // a 16- or 8-bit fallback scan needs the slices' per-shell maxima: queued
    // behind the copies, they run while the host applies the rule
    if (fallbackPrecision == THX_FP16 || fallbackPrecision == THX_FP8)
        THX_TRY(ensure_slice_shell_max(c));
    THX_HIP(hipEventSynchronize(c->lpRead));
    c->lpScore.assign(hScore, hScore + P);
    c->lpEssH.assign(hEss, hEss + P);
    std::vector<int> bR(hR, hR + P), bT(hT, hT + P);
    std::vector<float> spread(hSpread, hSpread + P), tail(hTail, hTail + P);
    c->lpBestR = bR;
    c->lpBestT = bT;
    c->lpFell.assign(P, 0);
    c->lpSpreadH.assign(spread.begin(), spread.end());
    c->lpTailH.assign(tail.begin(), tail.end());
    // the fallback rule: collapsed ESS, a diffuse proposal, a best pose in
    // the proposal's tail, or a best score far below the batch median
    double med = 0.0;
    if (scoreDrop > 0.0) {
        std::vector<double> sc(c->lpScore);
        std::nth_element(sc.begin(), sc.begin() + P / 2, sc.end());
        med = sc[P / 2];
    }
    std::vector<int> flagged;
    for (int p = 0; p < P; p++) {
        const bool fb = c->lpBestR[p] < 0 || c->lpBestT[p] < 0 ||
                        (essMin > 0.0 && !(c->lpEssH[p] >= essMin)) ||
                        (spreadMax > 0.0 && !(spread[p] <= spreadMax)) ||
                        (tailMax > 0.0 && !(tail[p] <= tailMax)) ||
                        (scoreDrop > 0.0 && !(c->lpScore[p] >= med - scoreDrop));
        if (fb) flagged.push_back(p);
    }
    // uniform whole-grid search of the flagged particles (tiled scan)
    const int nf = (int)flagged.size();
    if (nf > 0) {
        const int nev0 = c->nev;
        // the indices go up from the pinned staging (its read-back contents
        // were copied out above; the stream finishes this upload before the
        // fallback's results are read back into it)
        int* pinned = c->lpStage.as<int>();
        std::copy(flagged.begin(), flagged.end(), pinned);
        const void *srcS, *srcR, *srcT;
        if (c->lpWideR > 0) {
            // a wider selection of the same proposals instead of the whole grid
            // (its two stages are timed: fallback_select, fallback_score)
            THX_TRY(rescore_wide(c, C, nf, pinned));
            srcS = c->locBase.p;
            srcR = c->locBestR.p;
            srcT = c->locBestT.p;
        } else {
            THX_TRY(scan_subset(c, cls, fallbackPrecision, flagged, pinned));
            srcS = c->lpRes.bestScore.p;
            srcR = c->lpRes.bestR.p;
            srcT = c->lpRes.bestT.p;
            c->nev = nev0;   // keep the stage events of this call, not the scan's
            c->evNames.resize(nev0);
        }
        // the staging holds P >= nf entries of each kind
        double* fS = c->lpStage.as<double>();
        int* fR = reinterpret_cast<int*>(fS + nf);
        int* fT = fR + nf;
        THX_HIP(hipMemcpyAsync(fS, srcS, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
        THX_HIP(hipMemcpyAsync(fR, srcR, sizeof(int) * nf, hipMemcpyDeviceToHost, st));
        THX_HIP(hipMemcpyAsync(fT, srcT, sizeof(int) * nf, hipMemcpyDeviceToHost, st));
        THX_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < nf; i++) {
            const int p = flagged[i];
            c->lpBestR[p] = fR[i];
            c->lpBestT[p] = fT[i];
            c->lpScore[p] = fS[i];
            c->lpFell[p] = 1;
        }
    }
    THX_TRY(record(c, "fallback"));
    THX_HIP(hipStreamSynchronize(st));
    if (nFallback) *nFallback = nf;
    c->lpValid = true;
    c->locValid = false;   // loc* buffers now hold learned-mode data
    return THX_OK;
}

int thx_ctx_get_learned(thx_ctx* c, int* bestR, int* bestT, double* score, double* ess,
                        float* spread, float* tail, int* fellBack) {
    clear_error();
    THX_CHECK(c && c->lpValid, THX_ERR_STATE, "get_learned: no learned expectation done");
    const int P = (int)c->lpBestR.size();
    if (bestR) std::memcpy(bestR, c->lpBestR.data(), sizeof(int) * P);
    if (bestT) std::memcpy(bestT, c->lpBestT.data(), sizeof(int) * P);
    if (score) std::memcpy(score, c->lpScore.data(), sizeof(double) * P);
    if (ess) std::memcpy(ess, c->lpEssH.data(), sizeof(double) * P);
    for (int p = 0; p < P; p++) {
        if (spread) spread[p] = (float)c->lpSpreadH[p];
        if (tail) tail[p] = (float)c->lpTailH[p];
    }
    if (fellBack) std::memcpy(fellBack, c->lpFell.data(), sizeof(int) * P);
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Particle-filter entry points (host buffers; ragged counts per particle)
// ---------------------------------------------------------------------------

static std::vector<int> offsets(const int* cnt, int P, size_t& total) {
    std::vector<int> o(P);
    total = 0;
    for (int p = 0; p < P; p++) {
        o[p] = (int)total;
        total += cnt[p];
    }
    return o;
}

int thx_pf_resample(thx_ctx* c, const double* w, const double* u, const int* nIn,
                    const int* nOut, int P, uint64_t seed, const double* u0, int* idx,
                    double* wNew, double* ess) {
    clear_error();
    THX_CHECK(c && w && u && nIn && nOut && idx && wNew && P > 0, THX_ERR_ARG,
              "pf_resample: bad argument");
    for (int p = 0; p < P; p++)
        THX_CHECK(nIn[p] >= 1 && nIn[p] <= kResampleMaxN && nOut[p] >= 1, THX_ERR_RANGE,
                  "pf_resample: counts must be in [1, 8192]");
    THX_TRY(set_device(c));
    size_t tin = 0, tout = 0;
    std::vector<int> oin = offsets(nIn, P, tin), oout = offsets(nOut, P, tout);
    DevBuf dw, du, doi, dni, doo, dno, du0, didx, dwn, dess;
    int s = THX_OK;
    hipStream_t st = c->stream;
    if ((s = upload(dw, w, tin, st)) || (s = upload(du, u, tin, st)) ||
        (s = upload(doi, oin.data(), P, st)) || (s = upload(dni, nIn, P, st)) ||
        (s = upload(doo, oout.data(), P, st)) || (s = upload(dno, nOut, P, st)) ||
        (u0 && (s = upload(du0, u0, P, st))) || (s = didx.ensure(sizeof(int) * tout)) ||
        (s = dwn.ensure(sizeof(double) * tout)) || (s = dess.ensure(sizeof(double) * P)))
        return s;
    s = pf_resample(st, dw.as<double>(), du.as<double>(), doi.as<int>(), dni.as<int>(),
                    doo.as<int>(), dno.as<int>(), P, seed, 0, u0 ? du0.as<double>() : nullptr,
                    didx.as<int>(), dwn.as<double>(), dess.as<double>());
    if (s == THX_OK) {
        THX_HIP(hipStreamSynchronize(st));
        THX_HIP(hipMemcpy(idx, didx.p, sizeof(int) * tout, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(wNew, dwn.p, sizeof(double) * tout, hipMemcpyDeviceToHost));
        if (ess) THX_HIP(hipMemcpy(ess, dess.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    }
    DevBuf* all[] = {&dw, &du, &doi, &dni, &doo, &dno, &du0, &didx, &dwn, &dess};
    for (DevBuf* b : all) b->release();
    return s;
}

int thx_pf_rot_mean(thx_ctx* c, const double* quat, const double* w, const int* cnt, int P,
                    double* mean, double* k) {
    clear_error();
    THX_CHECK(c && quat && cnt && mean && k && P > 0, THX_ERR_ARG, "pf_rot_mean: bad argument");
    THX_TRY(set_device(c));
    size_t tot = 0;
    std::vector<int> off = offsets(cnt, P, tot);
    DevBuf dq, dw, doff, dcnt, dm, dk;
    hipStream_t st = c->stream;
    int s = THX_OK;
    if ((s = upload(dq, quat, 4 * tot, st)) || (w && (s = upload(dw, w, tot, st))) ||
        (s = upload(doff, off.data(), P, st)) || (s = upload(dcnt, cnt, P, st)) ||
        (s = dm.ensure(sizeof(double) * 4 * P)) || (s = dk.ensure(sizeof(double) * 3 * P)))
        return s;
    s = pf_rot_mean(st, dq.as<double>(), w ? dw.as<double>() : nullptr, doff.as<int>(),
                    dcnt.as<int>(), P, dm.as<double>(), dk.as<double>());
    if (s == THX_OK) {
        THX_HIP(hipStreamSynchronize(st));
        THX_HIP(hipMemcpy(mean, dm.p, sizeof(double) * 4 * P, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(k, dk.p, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    }
    DevBuf* all[] = {&dq, &dw, &doff, &dcnt, &dm, &dk};
    for (DevBuf* b : all) b->release();
    return s;
}

int thx_pf_perturb(thx_ctx* c, double* quat, const int* nR, double* trans, const int* nT, int P,
                   const double* mean, const double* k, const double* s01rho, double pf,
                   uint64_t seed, int phase, const double* d) {
    clear_error();
    THX_CHECK(c && P > 0 && ((quat && nR && mean && k) || (trans && nT && s01rho)), THX_ERR_ARG,
              "pf_perturb: bad argument");
    THX_TRY(set_device(c));
    hipStream_t st = c->stream;
    int s = THX_OK;
    if (quat) {
        size_t tot = 0;
        std::vector<int> off = offsets(nR, P, tot);
        int mx = 0;
        for (int p = 0; p < P; p++) mx = std::max(mx, nR[p]);
        DevBuf dq, doff, dcnt, dm, dk, dd;
        if ((s = upload(dq, quat, 4 * tot, st)) || (s = upload(doff, off.data(), P, st)) ||
            (s = upload(dcnt, nR, P, st)) || (s = upload(dm, mean, 4 * (size_t)P, st)) ||
            (s = upload(dk, k, 3 * (size_t)P, st)) || (d && (s = upload(dd, d, 4 * tot, st))))
            return s;
        s = pf_perturb_rot(st, dq.as<double>(), doff.as<int>(), dcnt.as<int>(), mx,
                           dm.as<double>(), dk.as<double>(), 0.0, pf, P, seed, (uint32_t)phase,
                           d ? dd.as<double>() : nullptr);
        if (s == THX_OK) {
            THX_HIP(hipStreamSynchronize(st));
            THX_HIP(hipMemcpy(quat, dq.p, sizeof(double) * 4 * tot, hipMemcpyDeviceToHost));
        }
        DevBuf* all[] = {&dq, &doff, &dcnt, &dm, &dk, &dd};
        for (DevBuf* b : all) b->release();
        if (s != THX_OK) return s;
    }
    if (trans) {
        size_t tot = 0;
        std::vector<int> off = offsets(nT, P, tot);
        int mx = 0;
        for (int p = 0; p < P; p++) mx = std::max(mx, nT[p]);
        DevBuf dt, doff, dcnt, ds;
        if ((s = upload(dt, trans, 2 * tot, st)) || (s = upload(doff, off.data(), P, st)) ||
            (s = upload(dcnt, nT, P, st)) || (s = upload(ds, s01rho, 3 * (size_t)P, st)))
            return s;
        s = pf_perturb_trans(st, dt.as<double>(), doff.as<int>(), dcnt.as<int>(), mx,
                             ds.as<double>(), pf, P, seed, (uint32_t)phase);
        if (s == THX_OK) {
            THX_HIP(hipStreamSynchronize(st));
            THX_HIP(hipMemcpy(trans, dt.p, sizeof(double) * 2 * tot, hipMemcpyDeviceToHost));
        }
        DevBuf* all[] = {&dt, &doff, &dcnt, &ds};
        for (DevBuf* b : all) b->release();
    }
    return s;
}

int thx_proposal_sample(thx_ctx* c, const double* mix, int C, int P, int nRot, int nTr,
                        uint64_t seed, double* quat, double* qdens, double* trans) {
    clear_error();
    THX_CHECK(c && mix && quat && qdens && trans && C > 0 && P > 0 && nRot > 0 && nTr > 0,
              THX_ERR_ARG, "proposal_sample: bad argument");
    THX_TRY(set_device(c));
    hipStream_t st = c->stream;
    DevBuf dm, dq, dd, dt;
    int s = THX_OK;
    if ((s = upload(dm, mix, (size_t)P * (6 * C + 3), st)) ||
        (s = dq.ensure(sizeof(double) * 4 * (size_t)P * nRot)) ||
        (s = dd.ensure(sizeof(double) * (size_t)P * nRot)) ||
        (s = dt.ensure(sizeof(double) * 2 * (size_t)P * nTr)))
        return s;
    s = pf_sample_mixture(st, dm.as<double>(), C, P, nRot, nTr, seed, 0, dq.as<double>(),
                          dd.as<double>(), dt.as<double>());
    if (s == THX_OK) {
        THX_HIP(hipStreamSynchronize(st));
        THX_HIP(hipMemcpy(quat, dq.p, sizeof(double) * 4 * (size_t)P * nRot, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(qdens, dd.p, sizeof(double) * (size_t)P * nRot, hipMemcpyDeviceToHost));
        THX_HIP(hipMemcpy(trans, dt.p, sizeof(double) * 2 * (size_t)P * nTr, hipMemcpyDeviceToHost));
    }
    DevBuf* all[] = {&dm, &dq, &dd, &dt};
    for (DevBuf* b : all) b->release();
    return s;
}

// ---------------------------------------------------------------------------
// Synthetic workloads
// ---------------------------------------------------------------------------

static void philox_normals(uint64_t seed, uint32_t a, uint32_t stream, double out[4]) {
    uint32_t w[4], v[4];
    Philox::draw(seed, a, 0, stream, w);
    Philox::draw(seed, a, 1, stream, v);
    const double u1 = Philox::u01(w[0], w[1]), u2 = Philox::u01(w[2], w[3]);
    const double u3 = Philox::u01(v[0], v[1]), u4 = Philox::u01(v[2], v[3]);
    const double r1 = std::sqrt(-2 * std::log(u1)), r2 = std::sqrt(-2 * std::log(u3));
    out[0] = r1 * std::cos(2 * M_PI * u2);
    out[1] = r1 * std::sin(2 * M_PI * u2);
    out[2] = r2 * std::cos(2 * M_PI * u4);
    out[3] = r2 * std::sin(2 * M_PI * u4);
}

static void philox_uniforms(uint64_t seed, uint32_t a, uint32_t stream, double out[2]) {
    uint32_t w[4];
    Philox::draw(seed, a, 2, stream, w);
    out[0] = Philox::u01(w[0], w[1]);
    out[1] = Philox::u01(w[2], w[3]);
}

int thx_ctx_synth_volume(thx_ctx* c, int cls, int idim, int pf, int nBlob, double sigma,
                         double rBall, uint64_t seed, double* centresOut) {
    clear_error();
    THX_CHECK(c && cls >= 0 && idim > 0 && idim % 2 == 0 && pf > 0 && nBlob > 0 && sigma > 0,
              THX_ERR_ARG, "synth_volume: bad argument");
    THX_TRY(set_device(c));
    std::vector<double> cen(3 * (size_t)nBlob);
    for (int b = 0; b < nBlob; b++) {
        double g[4], u[2];
        philox_normals(seed, (uint32_t)b, kStreamBlob, g);
        philox_uniforms(seed, (uint32_t)b, kStreamBlob, u);
        const double nrm = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double rad = rBall * std::cbrt(u[0]);
        for (int a = 0; a < 3; a++) cen[3 * b + a] = g[a] / nrm * rad;
    }
    if (centresOut) std::memcpy(centresOut, cen.data(), sizeof(double) * cen.size());
    return thx_ctx_synth_volume_blobs(c, cls, idim, pf, nBlob, sigma, cen.data(), nullptr);
}

int thx_ctx_synth_volume_blobs(thx_ctx* c, int cls, int idim, int pf, int nBlob, double sigma,
                               const double* centres, const double* amps) {
    clear_error();
    THX_CHECK(c && cls >= 0 && idim > 0 && idim % 2 == 0 && pf > 0 && nBlob > 0 && sigma > 0 &&
                  centres,
              THX_ERR_ARG, "synth_volume: bad argument");
    THX_TRY(set_device(c));
    std::vector<double> cen(centres, centres + 3 * (size_t)nBlob);
    if (amps) cen.insert(cen.end(), amps, amps + nBlob);
    const int V = idim * pf;
    if ((int)c->vol.size() <= cls) {
        c->vol.resize(cls + 1);
        c->vdim.resize(cls + 1, 0);
    }
    THX_TRY(c->vol[cls].ensure(sizeof(float2) * (size_t)(V / 2 + 1) * V * V));
    c->vdim[cls] = V;
    invalidate_vol_lo(c, cls);
    DevBuf dc;
    THX_TRY(dc.ensure(sizeof(double) * cen.size()));
    THX_HIP(hipMemcpy(dc.p, cen.data(), sizeof(double) * cen.size(), hipMemcpyHostToDevice));
    int s = launch_synth_volume(c, cls, idim, pf, dc.as<const double>(), nBlob, sigma,
                                amps ? dc.as<const double>() + 3 * (size_t)nBlob : nullptr);
    hipError_t e = hipStreamSynchronize(c->stream);
    dc.release();
    if (s != THX_OK) return s;
    THX_CHECK(e == hipSuccess, THX_ERR_HIP, "synth_volume: kernel failed");
    return THX_OK;
}

int thx_ctx_synth_particles(thx_ctx* c, int cls, int P, double maxShift, double dfMin,
                            double dfMax, double voltage, double Cs, double ampContrast,
                            double pixelSize, double snr, uint64_t seed, double* trueQuat,
                            double* trueShift, double* ctfParams) {
    clear_error();
    THX_CHECK(c && P > 0 && snr > 0 && pixelSize > 0, THX_ERR_ARG, "synth_particles: bad argument");
    THX_CHECK(c->K > 0, THX_ERR_STATE, "synth_particles: pixel set not defined");
    THX_CHECK(cls >= 0 && cls < (int)c->vol.size() && c->vdim[cls] == c->idim * c->pf,
              THX_ERR_STATE, "synth_particles: volume of this class not set");
    THX_TRY(set_device(c));
    std::vector<double> prm(9 * (size_t)P);
    for (int p = 0; p < P; p++) {
        double g[4], u[2], v[2];
        philox_normals(seed, (uint32_t)p, kStreamPose, g);
        const double n = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
        for (int a = 0; a < 4; a++) prm[9 * (size_t)p + a] = g[a] / n;
        philox_uniforms(seed, (uint32_t)p, kStreamShift, u);
        prm[9 * (size_t)p + 4] = (2 * u[0] - 1) * maxShift;
        prm[9 * (size_t)p + 5] = (2 * u[1] - 1) * maxShift;
        philox_uniforms(seed, (uint32_t)p, kStreamCtf, v);
        double w[2];
        philox_uniforms(seed ^ 0x9e3779b97f4a7c15ull, (uint32_t)p, kStreamCtf, w);
        prm[9 * (size_t)p + 6] = dfMin + (dfMax - dfMin) * v[0];
        prm[9 * (size_t)p + 7] = dfMin + (dfMax - dfMin) * v[1];
        prm[9 * (size_t)p + 8] = M_PI * w[0];
        if (trueQuat)
            for (int a = 0; a < 4; a++) trueQuat[4 * (size_t)p + a] = prm[9 * (size_t)p + a];
        if (trueShift) {
            trueShift[2 * (size_t)p] = prm[9 * (size_t)p + 4];
            trueShift[2 * (size_t)p + 1] = prm[9 * (size_t)p + 5];
        }
        if (ctfParams)
            for (int a = 0; a < 3; a++) ctfParams[3 * (size_t)p + a] = prm[9 * (size_t)p + 6 + a];
    }
    THX_TRY(alloc_particles(c, P));
    DevBuf dp, de;
    THX_TRY(dp.ensure(sizeof(double) * prm.size()));
    const size_t nbx = (size_t)(c->K + 255) / 256;   // signal-energy partials per particle
    THX_TRY(de.ensure(sizeof(double) * P * nbx));
    THX_HIP(hipMemcpy(dp.p, prm.data(), sizeof(double) * prm.size(), hipMemcpyHostToDevice));
    THX_TRY(launch_synth_signal(c, cls, P, dp.as<const double>(), de.as<double>(), voltage, Cs,
                                ampContrast, pixelSize));
    std::vector<double> part((size_t)P * nbx), en(P, 0.0);
    THX_HIP(hipStreamSynchronize(c->stream));
    THX_HIP(hipMemcpy(part.data(), de.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < P; p++)
        for (size_t b = 0; b < nbx; b++) en[p] += part[(size_t)p * nbx + b];
    dp.release();
    de.release();
    // real-space variance of the band-limited signal (Parseval over the
    // half plane, doubled): var = 2 sum |S_k|^2 / N^4; sigma_n^2 = var / SNR.
    const double N = c->idim;
    double mean = 0;
    for (int p = 0; p < P; p++) mean += 2 * en[p] / (N * N * N * N);
    mean /= P;
    const double sn2 = mean / snr;
    // per-component sd of the unnormalised DFT of real white noise: N sigma_n / sqrt(2)
    const double compSigma = N * std::sqrt(sn2 / 2);
    THX_TRY(launch_synth_noise(c, P, compSigma, seed ^ 0x6e6f697365ull));
    THX_HIP(hipStreamSynchronize(c->stream));
    c->ctfReady = c->sigReady = true;
    return THX_OK;
}

// ---------------------------------------------------------------------------
// Reference-shaped entry points
// ---------------------------------------------------------------------------

__global__ void k_rotmat_colmajor(const double* __restrict__ q, int nR, double* __restrict__ m) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nR) return;
    const double w = q[4 * r], x = q[4 * r + 1], y = q[4 * r + 2], z = q[4 * r + 3];
    const double A[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const double aa = A[a][0] * A[0][b] + A[a][1] * A[1][b] + A[a][2] * A[2][b];
            m[9 * (size_t)r + a + 3 * b] = (a == b ? 1.0 : 0.0) + 2 * w * A[a][b] + 2 * aa;
        }
}

__global__ __launch_bounds__(256) void k_project_mat(
    const float2* __restrict__ vol, int V, int pf, const double* __restrict__ mat, int R,
    const int* __restrict__ iCol, const int* __restrict__ iRow, int K, float2* __restrict__ out) {
    // grid (R, ceil(K / 256)): rotations on x, the dimension with the largest limit
    const int r = blockIdx.x;
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= K || r >= R) return;
    const double* m = mat + 9 * (size_t)r;   // column-major
    const double c0 = (double)(iCol[k] * pf), c1 = (double)(iRow[k] * pf);
    const float x = (float)(m[0] * c0 + m[3] * c1);
    const float y = (float)(m[1] * c0 + m[4] * c1);
    const float z = (float)(m[2] * c0 + m[5] * c1);
    bool conj = false;
    float xx = x, yy = y, zz = z;
    if (xx < 0) { xx = -xx; yy = -yy; zz = -zz; conj = true; }
    const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
    const float xd = xx - x0, yd = yy - y0, zd = zz - z0;
    const float v0[2] = {1 - xd, xd}, v1[2] = {1 - yd, yd}, v2[2] = {1 - zd, zd};
    const int nColFT = V / 2 + 1;
    float re = 0.f, im = 0.f;
    for (int kk = 0; kk < 2; kk++) {
        int zi = z0 + kk;
        zi = zi < 0 ? zi + V : zi;
        for (int j = 0; j < 2; j++) {
            int yi = y0 + j;
            yi = yi < 0 ? yi + V : yi;
            for (int i = 0; i < 2; i++) {
                const float w = v0[i] * v1[j] * v2[kk];
                const float2 d = vol[((size_t)zi * V + yi) * nColFT + x0 + i];
                re += d.x * w;
                im += d.y * w;
            }
        }
    }
    out[(size_t)r * K + k] = conj ? make_float2(re, -im) : make_float2(re, im);
}

__global__ void k_norm_rows(const float2* __restrict__ rotP, int n, float* __restrict__ normP) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n) return;
    const float2 v = rotP[i];
    normP[i] = v.x * v.x + v.y * v.y;
}

int thx_ExpectRotran(float* traP, const double* trans, const double* rot, double* rotMat,
                     const int* iCol, const int* iRow, int nR, int nT, int idim, int npxl) {
    clear_error();
    THX_CHECK(traP && trans && rot && rotMat && iCol && iRow, THX_ERR_ARG,
              "ExpectRotran: null argument");
    THX_CHECK(nR > 0 && nT > 0 && idim > 0 && npxl > 0, THX_ERR_ARG, "ExpectRotran: bad size");
    thx_ctx* c = nullptr;
    THX_TRY(default_ctx(&c));
    THX_TRY(set_device(c));
    DevBuf dCol, dRow, dTr, dS, dQ, dM;
    int s = THX_OK;
    do {
        if ((s = dCol.ensure(sizeof(int) * npxl)) || (s = dRow.ensure(sizeof(int) * npxl)) ||
            (s = dTr.ensure(sizeof(double) * 2 * nT)) ||
            (s = dS.ensure(sizeof(float2) * (size_t)nT * npxl)) ||
            (s = dQ.ensure(sizeof(double) * 4 * nR)) || (s = dM.ensure(sizeof(double) * 9 * nR)))
            break;
        if (hipMemcpy(dCol.p, iCol, sizeof(int) * npxl, hipMemcpyHostToDevice) ||
            hipMemcpy(dRow.p, iRow, sizeof(int) * npxl, hipMemcpyHostToDevice) ||
            hipMemcpy(dTr.p, trans, sizeof(double) * 2 * nT, hipMemcpyHostToDevice) ||
            hipMemcpy(dQ.p, rot, sizeof(double) * 4 * nR, hipMemcpyHostToDevice)) {
            set_error("ExpectRotran: upload failed");
            s = THX_ERR_HIP;
            break;
        }
        if ((s = launch_translate_raw(c->stream, dTr.as<const double>(), nT, nT,
                                      dCol.as<const int>(), dRow.as<const int>(), npxl, npxl,
                                      idim, dS.as<float2>())))
            break;
        hipLaunchKernelGGL(k_rotmat_colmajor, dim3((nR + 127) / 128), dim3(128), 0, c->stream,
                           dQ.as<const double>(), nR, dM.as<double>());
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(traP, dS.p, sizeof(float2) * (size_t)nT * npxl, hipMemcpyDeviceToHost) ||
            hipMemcpy(rotMat, dM.p, sizeof(double) * 9 * nR, hipMemcpyDeviceToHost)) {
            set_error("ExpectRotran: kernel or download failed");
            s = THX_ERR_HIP;
        }
    } while (0);
    dCol.release(); dRow.release(); dTr.release(); dS.release(); dQ.release(); dM.release();
    return s;
}

int thx_ExpectProject(const float* volume, float* rotP, const double* rotMat, const int* iCol,
                      const int* iRow, int nR, int pf, int interp, int vdim, int npxl) {
    clear_error();
    THX_CHECK(volume && rotP && rotMat && iCol && iRow, THX_ERR_ARG,
              "ExpectProject: null argument");
    THX_CHECK(nR > 0 && pf > 0 && vdim > 0 && vdim % 2 == 0 && npxl > 0, THX_ERR_ARG,
              "ExpectProject: bad size");
    THX_CHECK(interp == 1, THX_ERR_ARG, "ExpectProject: only LINEAR_INTERP (1) is supported");
    double maxR = 0;
    for (int k = 0; k < npxl; k++) {
        THX_CHECK(iCol[k] >= 0, THX_ERR_ARG, "ExpectProject: iCol must be >= 0");
        maxR = std::max(maxR, std::hypot((double)iCol[k], (double)iRow[k]));
    }
    THX_CHECK(maxR * pf + 2 <= vdim / 2, THX_ERR_RANGE,
              "ExpectProject: pixel radius too close to Nyquist of the padded volume");
    thx_ctx* c = nullptr;
    THX_TRY(default_ctx(&c));
    THX_TRY(set_device(c));
    const size_t nv = (size_t)(vdim / 2 + 1) * vdim * vdim;
    DevBuf dV, dCol, dRow, dM, dOut;
    int s = THX_OK;
    do {
        if ((s = dV.ensure(sizeof(float2) * nv)) || (s = dCol.ensure(sizeof(int) * npxl)) ||
            (s = dRow.ensure(sizeof(int) * npxl)) || (s = dM.ensure(sizeof(double) * 9 * nR)) ||
            (s = dOut.ensure(sizeof(float2) * (size_t)nR * npxl)))
            break;
        if (hipMemcpy(dV.p, volume, sizeof(float2) * nv, hipMemcpyHostToDevice) ||
            hipMemcpy(dCol.p, iCol, sizeof(int) * npxl, hipMemcpyHostToDevice) ||
            hipMemcpy(dRow.p, iRow, sizeof(int) * npxl, hipMemcpyHostToDevice) ||
            hipMemcpy(dM.p, rotMat, sizeof(double) * 9 * nR, hipMemcpyHostToDevice)) {
            set_error("ExpectProject: upload failed");
            s = THX_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(k_project_mat, dim3(nR, (npxl + 255) / 256), dim3(256), 0, c->stream,
                           dV.as<const float2>(), vdim, pf, dM.as<const double>(), nR,
                           dCol.as<const int>(), dRow.as<const int>(), npxl, dOut.as<float2>());
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(rotP, dOut.p, sizeof(float2) * (size_t)nR * npxl, hipMemcpyDeviceToHost)) {
            set_error("ExpectProject: kernel or download failed");
            s = THX_ERR_HIP;
        }
    } while (0);
    dV.release(); dCol.release(); dRow.release(); dM.release(); dOut.release();
    return s;
}

int thx_ctx_project(thx_ctx* c, int cls, int pf, const double* rotMat, int nR, const int* iCol,
                    const int* iRow, int nPxl, float* out) {
    clear_error();
    THX_CHECK(c && rotMat && iCol && iRow && out && nR > 0 && nPxl > 0 && pf > 0, THX_ERR_ARG,
              "ctx_project: bad argument");
    THX_CHECK(cls >= 0 && cls < (int)c->vol.size() && c->vol[cls].p, THX_ERR_STATE,
              "ctx_project: no volume for this class");
    const int vdim = c->vdim[cls];
    double maxR = 0;
    for (int k = 0; k < nPxl; k++) {
        THX_CHECK(iCol[k] >= 0, THX_ERR_ARG, "ctx_project: iCol must be >= 0");
        maxR = std::max(maxR, std::hypot((double)iCol[k], (double)iRow[k]));
    }
    THX_CHECK(maxR * pf + 2 <= vdim / 2, THX_ERR_RANGE,
              "ctx_project: pixel radius too close to Nyquist of the padded volume");
    THX_TRY(set_device(c));
    hipStream_t st = c->stream;
    DevBuf dCol, dRow, dM, dOut;
    int s = THX_OK;
    if ((s = upload(dCol, iCol, nPxl, st)) || (s = upload(dRow, iRow, nPxl, st)) ||
        (s = upload(dM, rotMat, 9 * (size_t)nR, st)) ||
        (s = dOut.ensure(sizeof(float2) * (size_t)nR * nPxl)))
        return s;
    hipLaunchKernelGGL(k_project_mat, dim3(nR, (nPxl + 255) / 256), dim3(256), 0, st,
                       c->vol[cls].as<const float2>(), vdim, pf, dM.as<const double>(), nR,
                       dCol.as<const int>(), dRow.as<const int>(), nPxl, dOut.as<float2>());
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(out, dOut.p, sizeof(float2) * (size_t)nR * nPxl, hipMemcpyDeviceToHost) !=
            hipSuccess) {
        set_error("ctx_project: kernel or download failed");
        s = THX_ERR_HIP;
    }
    DevBuf* all[] = {&dCol, &dRow, &dM, &dOut};
    for (DevBuf* b : all) b->release();
    return s;
}

int thx_ExpectGlobal3D(const float* rotP, const float* traP, const float* datP, const float* ctfP,
                       const float* sigRcpP, float* wC, float* wR, float* wT, const double* pR,
                       const double* pT, float* baseL, int kIdx, int nK, int nR, int nT, int npxl,
                       int imgNum) {
    clear_error();
    THX_CHECK(rotP && traP && datP && ctfP && sigRcpP && wC && wR && wT && pR && pT && baseL,
              THX_ERR_ARG, "ExpectGlobal3D: null argument");
    THX_CHECK(nK > 0 && kIdx >= 0 && kIdx < nK && nR > 0 && nT > 0 && npxl > 0 && imgNum > 0,
              THX_ERR_ARG, "ExpectGlobal3D: bad size");
    thx_ctx* c = nullptr;
    THX_TRY(default_ctx(&c));
    THX_TRY(set_device(c));
    // pixel set of npxl entries (shell indices are not needed at FP32)
    c->K = npxl;
    c->Kp = round_up(npxl, kKAlign);
    c->maxShell = 0;
    THX_TRY(c->iSig.ensure(sizeof(int) * c->Kp));
    THX_HIP(hipMemset(c->iSig.p, 0, sizeof(int) * c->Kp));
    THX_TRY(alloc_particles(c, imgNum));
    const size_t n = (size_t)imgNum * npxl;
    THX_HIP(hipMemcpy(c->dat.p, datP, sizeof(float2) * n, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->ctf.p, ctfP, sizeof(float) * n, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy(c->sig.p, sigRcpP, sizeof(float) * n, hipMemcpyHostToDevice));
    c->ctfReady = c->sigReady = true;
    THX_TRY(set_grid_shape(c, nR, nT));
    THX_TRY(upload_priors(c, pR, pT));
    // projections and phases into the padded layouts (zero padding)
    THX_HIP(hipMemset(c->rotP.p, 0, sizeof(float2) * (size_t)c->Rpad * c->Kp));
    c->rotLoPrec = -1;
    c->shellMaxPOk = false;
    THX_HIP(hipMemset(c->S.p, 0, sizeof(float2) * (size_t)c->TpadS * c->Kp));
    THX_HIP(hipMemcpy2D(c->rotP.p, sizeof(float2) * c->Kp, rotP, sizeof(float2) * npxl,
                        sizeof(float2) * npxl, nR, hipMemcpyHostToDevice));
    THX_HIP(hipMemcpy2D(c->S.p, sizeof(float2) * c->Kp, traP, sizeof(float2) * npxl,
                        sizeof(float2) * npxl, nT, hipMemcpyHostToDevice));
    const size_t nrp = (size_t)c->Rpad * c->Kp;
    hipLaunchKernelGGL(k_norm_rows, dim3((unsigned)((nrp + 255) / 256)), dim3(256), 0, c->stream,
                       c->rotP.as<const float2>(), (int)nrp, c->normP.as<float>());
    THX_HIP(hipGetLastError());
    c->nev = 0;
    c->evNames.clear();
    THX_TRY(record(c, "start"));
    THX_TRY(run_scan(c, 0, THX_FP32));
    THX_HIP(hipStreamSynchronize(c->stream));
    const int P = imgNum;
    std::vector<double> dwC(P), dwR((size_t)P * nR), dwT((size_t)P * nT), dB(P);
    const ClassResult& r = c->res[0];
    THX_HIP(hipMemcpy(dwC.data(), r.wC.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(dwR.data(), r.wR.p, sizeof(double) * P * nR, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(dwT.data(), r.wT.p, sizeof(double) * P * nT, hipMemcpyDeviceToHost));
    THX_HIP(hipMemcpy(dB.data(), r.base.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    // class merge with the running baseline (kernel_getMaxBase /
    // kernel_setBaseLine / kernel_UpdateW semantics)
    for (int l = 0; l < P; l++) {
        double base = dB[l];
        if (kIdx > 0) {
            const double old = baseL[l];
            if (dB[l] > old) {
                const double nf = std::exp(old - dB[l]);
                for (int k = 0; k < nK; k++) {
                    wC[(size_t)l * nK + k] = (float)(wC[(size_t)l * nK + k] * nf);
                    for (int i = 0; i < nR; i++)
                        wR[((size_t)l * nK + k) * nR + i] = (float)(wR[((size_t)l * nK + k) * nR + i] * nf);
                    for (int i = 0; i < nT; i++)
                        wT[((size_t)l * nK + k) * nT + i] = (float)(wT[((size_t)l * nK + k) * nT + i] * nf);
                }
                base = dB[l];
            } else {
                base = old;
            }
        }
        const double g = std::exp(dB[l] - base);
        wC[(size_t)l * nK + kIdx] = (float)(g * dwC[l]);
        for (int i = 0; i < nR; i++)
            wR[((size_t)l * nK + kIdx) * nR + i] += (float)(g * dwR[(size_t)l * nR + i]);
        for (int i = 0; i < nT; i++)
            wT[((size_t)l * nK + kIdx) * nT + i] += (float)(g * dwT[(size_t)l * nT + i]);
        baseL[l] = (float)base;
    }
    return THX_OK;
}

}  // extern "C"
