// Internal declarations shared by the HIP translation units of the
// expectation-step library.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "thunder_expect.h"

namespace thx {

void set_error(const std::string& msg);
void clear_error();

#define THX_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            thx::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_)); \
            return THX_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define THX_CHECK(cond, code, msg)      \
    do {                                \
        if (!(cond)) {                  \
            thx::set_error(msg);        \
            return (code);              \
        }                               \
    } while (0)

#define THX_TRY(expr)               \
    do {                            \
        int s_ = (expr);            \
        if (s_ != THX_OK) return s_; \
    } while (0)

// Device buffer that grows on demand and is freed with the context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Pinned host staging (hipHostMalloc), grown as needed: device-to-host
// copies into it are asynchronous and can share one stream synchronisation.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Host array -> device buffer (grown as needed), asynchronous on st.
template <class T>
int upload(DevBuf& d, const T* h, size_t n, hipStream_t st) {
    int s = d.ensure(sizeof(T) * (n ? n : 1));
    if (s) return s;
    if (n && hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error("upload: hipMemcpyAsync failed");
        return 2;
    }
    return 0;
}

// Tile geometry of the scoring kernel: a block covers 128 (particle,
// translation) rows x 128 rotations.  TC rows belong to one particle.
constexpr int kTileM = 128;
constexpr int kTileN = 128;
constexpr int kChunk = 32;     // complex pixels per K step
// padded pixel count Kp: a multiple of two K steps (zero pixels past K), so
// the whole-grid scorer's loop, unrolled by two, has no odd tail
constexpr int kKAlign = 2 * kChunk;
constexpr int kMaxTiles = 8192; // nTc * nRb per particle handled by the merge

struct ClassResult {
    DevBuf wC, wR, wT, base;   // double: P, P*R, P*T, P
    DevBuf bestR, bestT;       // int: P
    DevBuf bestScore;          // double: P
    bool valid = false;
};

// Device pointers of one ragged local-search batch.
struct LocalBatch {
    const float2* const* vols = nullptr;   // per-class volume base pointers
    const uint32_t* const* volsLo = nullptr;   // per-class BF16 copies (16-bit scorers)
    const int* pcls = nullptr;             // class per particle (or null: class 0)
    const double* quat = nullptr;          // [Rtot][4]
    const float* pRf = nullptr;            // [Rtot]
    const int* rOff = nullptr;             // [P]
    const int* nR = nullptr;               // [P]
    const double* trans = nullptr;         // [Ttot][2]
    const float* pTf = nullptr;            // [Ttot]
    const int* tOff = nullptr;             // [P]
    const int* nT = nullptr;               // [P]
    const int2* blocks = nullptr;          // [nblk] (particle, first rotation)
    const int* bStart = nullptr;           // [P + 1]
    float* bMax = nullptr;                 // [nblk]
    float* bWC = nullptr;                  // [nblk]
    int* bArg = nullptr;                   // [nblk]
    float* bWT = nullptr;                  // [nblk][kLocalWtStride] (selection scorer: [nblk][TPAD])
    float* wRrel = nullptr;                // [Rtot]
    // equal counts for every particle (rOff / nR / tOff / nT / bStart null):
    // particle p has affR rotations from p * affR, affT translations from
    // p * affT and blocks from p * affB
    int affR = 0, affT = 0, affB = 0;
};

struct Timing {
    std::vector<std::string> names;
    std::vector<double> ms;
};

}  // namespace thx

struct thx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // pixel set
    int K = 0, Kp = 0, idim = 0, pf = 2, maxShell = 0;
    thx::DevBuf iCol, iRow, iSig;
    // volumes per class, and their bricked BF16 copies for the 16-bit local
    // scorers (built on first use after a volume changes)
    std::vector<thx::DevBuf> vol;
    std::vector<int> vdim;
    std::vector<thx::DevBuf> volLo;
    std::vector<char> volLoOk;
    thx::DevBuf locVolsLo;
    thx::DevBuf lpQf;        // grid quaternions as float4 for the selection
    thx::DevBuf S16;         // FP16 copy of the phase table S (selected-pose scorer)
    thx::HostBuf lpStage;    // pinned: per-particle results read back in one batch
    thx::DevBuf lpStageD;    // the same batch packed on the device
    // ordering events between the context stream and the proposal network's
    hipEvent_t lpReady = nullptr, lpDone = nullptr, lpRead = nullptr;
    // projectee preparation: FFT plans (hipfftHandle) for one (N, pf), scratch
    void* prjPlanBw = nullptr;
    void* prjPlanFw = nullptr;
    int prjN = 0, prjPf = 0;
    thx::DevBuf prjIn, prjReal, prjPad;
    // particle images: batched R2C FFT plan (hipfftHandle) for (imgN, imgP), buffers
    void* imgPlan = nullptr;
    int imgN = 0, imgP = 0, imgB = 0;   // imgB transforms: P, or 2 P with the soft mask
    thx::DevBuf imgIn, imgReal, imgFT, imgMask;
    thx::DevBuf imgBgStat, imgScale;  // per image background sum of squares and count; stdN, scale
    bool imgReady = false;            // imgFT holds the spectra of the last images
    bool partFromImages = false;      // the particle arrays were packed from those images
    bool ctfReady = false, sigReady = false;   // the particles' ctf / sigRcp arrays are set
    thx::DevBuf ctfAttr;              // float [P][7]: CTF attributes of the last thx_ctx_set_ctf
    // reconstruction accumulators per class: F (double2) and T (double) on the
    // (V/2+1) V V half grid, and per class the insertDir sums and sample count
    std::vector<thx::DevBuf> recoF, recoT;
    std::vector<double> recoDir;
    int recoV = 0, recoCls = 0;
    bool ctfAttrOk = false;           // ctfAttr belongs to the current particles
    thx::DevBuf sigPart, sigShell;    // initial sigma: per-pixel partial sums, shell results
    int pixMaxCol = 0, pixMaxAbsRow = 0;   // extent of the pixel set
    // particles (image-major, as uploaded) and derived operands
    int P = 0, Ppad = 0;
    thx::DevBuf dat, ctf, sig;
    thx::DevBuf Aterm;      // double [Ppad]
    thx::DevBuf Y;          // float2 [Ppad][Kp]  (sig ctf conj(X)) * shell scale
    thx::DevBuf Wt;         // float  [Ppad][Kp]  sig ctf^2
    int prepPrecision = -1;
    // pose grid
    int R = 0, T = 0, Rpad = 0, TpadS = 0;  // Rpad: multiple of 256; TpadS: rows of S, pTf
    // shape of the current scan (chosen per precision by choose_scan_shape)
    int TC = 128, nTc = 1, Tpad = 0, NB = 128, nRb = 0;
    thx::DevBuf quat, trans, pR, pT, pRf, pTf;
    thx::DevBuf S;          // float2 [Tpad][Kp]
    // scan workspaces
    thx::DevBuf rotP;       // float2 [Rpad][Kp]
    thx::DevBuf rotLo;      // 2 x 16-bit (or 2 x 8-bit) [Rpad][Kp]
    int rotLoPrec = -1;     // precision rotLo holds for the current rotP; -1: stale
    thx::DevBuf normP;      // float [Rpad][Kp]  |P|^2
    thx::DevBuf Cterm;      // float [Ppad][Rpad]
    thx::DevBuf normPart;   // float [ns][Ppad][Rpad]: pixel-range slabs of Cterm
    thx::DevBuf shellScale; // float [maxShell+1]
    thx::DevBuf shellMax;   // float [2][maxShell+1] (scratch of the scale pass)
    bool shellMaxPOk = false;   // shellMax[0..maxShell] holds the current slices' maxima
    thx::DevBuf partMax, partWC, partArg, partWR, partWT;
    std::vector<thx::ClassResult> res;
    // ragged local-search batch (inputs, partials, results)
    int locP = 0, locTPAD = 0, locNblk = 0;
    size_t locRtot = 0, locTtot = 0;
    // rotation order of the last thx_ctx_score_local batch: scorer position ->
    // caller's index (empty: identity), and the per-particle offsets
    std::vector<int> locPerm, locROffH;
    thx::DevBuf locQuat, locTrans, locPRf, locPTf, locROff, locNR, locTOff, locNT, locBlocks,
        locBStart, locPcls, locVols, locBMax, locBWC, locBArg, locBWT, locWRrel;
    thx::DevBuf locWC, locWR, locWT, locBase, locBestR, locBestT;
    bool locValid = false;
    // launch shape of the local scorer: requested (0 = autotuned), the tuned
    // choice per (precision, translation padding, work-item count bucket),
    // the last one used, and the captured graph of a local phase
    int locModeReq = 0, locModeUsed = 0;   // launch form requested / last used
    int locNsReq = 0;                       // gather depth requested (0: tuned)
    bool locGraphOn = false;
    std::vector<std::pair<long long, int>> locTuned;
    thx::DevBuf locCounter;
    hipGraphExec_t locGraph = nullptr;
    std::vector<long long> locGraphKey;
    // 64 x 64 particle images of the learned proposal
    thx::DevBuf img;
    // grid-restricted learned proposal: mixture, selections, results
    thx::DevBuf lpMix, lpRidx, lpRlogq, lpTidx, lpESS, lpSpread, lpTail, lpPmap;
    // selected-pose scorer's particle order (Morton keys, sort scratch)
    thx::DevBuf lpOrder, lpOrdKeyA, lpOrdKeyB, lpOrdVal, lpOrdTmp;
    bool selOrder = true;
    // fallback of the learned expectation by a wider selection (0: whole grid)
    int lpWideR = 0, lpWideT = 0;
    thx::DevBuf lpMixSub;    // mixture rows of the flagged particles
    thx::DevBuf lpDat, lpCtf, lpSig;      // gathered rows of fallen-back particles
    thx::ClassResult lpRes;               // their whole-grid scan
    std::vector<int> lpBestR, lpBestT, lpFell;
    std::vector<double> lpScore, lpEssH, lpSpreadH, lpTailH;
    bool lpValid = false;
    // particle-filter state of the local search: mLR rotations and mLT shifts
    // per particle (FP64), weights, concentrations, deviations, tops
    int pfP = 0, pfR = 0, pfT = 0, pfLastPhase = 0;
    bool pfValid = false;
    thx::DevBuf pfQ, pfQ2, pfTr, pfTr2, pfWR, pfWT, pfWRf, pfWTf, pfIdx, pfWNew, pfK, pfS,
        pfMean, pfTopR, pfTopT, pfPeak, pfTransS, pfBest, pfActive, pfNPhase, pfOffR, pfOffT,
        pfNR, pfNT, pfNRa, pfNTa;
    // device work list of the phases: particle order, phase counter and counts
    // ([0] phase, [1] work items, [2] active at the start, [3] active at the
    // end), their pinned read-back, and the captured phase graphs (one per
    // assignment of the sample buffers, which swap every phase)
    thx::DevBuf pfOrder, pfCounts;
    thx::HostBuf pfCountsH;
    std::vector<std::pair<std::vector<long long>, hipGraphExec_t>> pfGraphs;
    // timing
    hipEvent_t ev[16] = {};
    int nev = 0;
    thx::Timing timing;
    std::vector<std::string> evNames;
};

namespace thx {
// Kernel launchers (thx_kernels.hip).  All enqueue on ctx->stream.
// outLo (optional, with out == rotP): the BF16 table written in the same pass
int launch_project(thx_ctx* c, int cls, const double* dQuat, int R, int Rpad,
                   float2* out, float* outNorm, uint32_t* outLo = nullptr);
int launch_translate_table(thx_ctx* c);
// S16 <- S in FP16 (TpadS x Kp), on the context stream
int launch_phase_half(thx_ctx* c);
int launch_translate_raw(hipStream_t st, const double* trans, int T, int Tpad, const int* iCol,
                         const int* iRow, int K, int Kp, int N, float2* S);
int launch_prep_particles(thx_ctx* c, int precision);
int launch_shell_scale(thx_ctx* c, int precision);
int ensure_slice_shell_max(thx_ctx* c);
int launch_convert_lo(thx_ctx* c, int precision);
int launch_norm_gemm(thx_ctx* c, int precision);
int launch_score(thx_ctx* c, int precision);
int score_tile_cols(int precision);
int launch_merge(thx_ctx* c, int cls);
int launch_synth_volume(thx_ctx* c, int cls, int idim, int pf,
                        const double* dCentres, int nBlob, double sigma,
                        const double* dAmps = nullptr);
int launch_synth_signal(thx_ctx* c, int cls, int P, const double* dParams,
                        double* dEnergy, double voltage, double Cs,
                        double ampContrast, double pixelSize);
int launch_synth_noise(thx_ctx* c, int P, double noiseSigma, uint64_t seed);
int launch_ctf_attr(thx_ctx* c, const float* dAttr, int P, float pixelSize);
int launch_score_one(thx_ctx* c, int cls, int p, float* dOut);
constexpr int kLocalRotBlock = 128;   // rotations per local-search block
// translation tiles of the local scorer (rows of a block's A operand) and the
// stride of its per-item translation partials (the largest tile)
constexpr int kLocalTiles = 3;
constexpr int kLocalTile[kLocalTiles] = {32, 64, 128};
constexpr int kLocalWtStride = 128;
inline int local_tile_index(int T) { return T <= 32 ? 0 : (T <= 64 ? 1 : 2); }
// the work items from `start` on (per-item outputs offset with them)
inline LocalBatch local_batch_range(const LocalBatch& b, int start) {
    LocalBatch r = b;
    r.blocks = b.blocks + start;
    r.bMax = b.bMax + start;
    r.bWC = b.bWC + start;
    r.bArg = b.bArg + start;
    r.bWT = b.bWT + (size_t)start * kLocalWtStride;
    return r;
}
// local scorer launch shapes: one block per work item, or a persistent grid
// pulling items from a counter
constexpr int kLocalDirect = 1;
constexpr int kLocalQueue = 2;
// A launch shape of the local scorer: the launch form (direct / persistent)
// and the gather depth ns, the trilinear samples per step whose corner loads
// are all in flight together (4 or 8 in the 16-bit forms; 0 means 4).
inline int local_shape(int mode, int ns) { return mode | (ns << 4); }
inline int local_shape_mode(int shape) { return shape & 15; }
inline int local_shape_ns(int shape) { return (shape >> 4) == 8 ? 8 : 4; }
// nItemsDev: the item count in device memory (nblk then bounds it)
int launch_score_local(thx_ctx* c, int precision, int TPAD, int nblk, const LocalBatch& b,
                       int shape = kLocalDirect, const int* nItemsDev = nullptr);
// particle order for the local scorer's work list (pose locality)
std::vector<int> pose_order(const double* quat, const int* off, const int* cls, int P,
                            const int* cnt = nullptr);
uint32_t quat_morton(const double* q);
// proposal network on its own stream after `ready`, `done` recorded at the end
int cnn_propose_async(thx_cnn* n, thx_ctx* c, double* mixDev, hipEvent_t ready,
                      hipEvent_t done);
// FP16 shell scales of the local scorer from the used class volumes + particles
int launch_local_shell_scale(thx_ctx* c, const std::vector<char>& used);
// BF16 copies of the used class volumes (built where missing or stale)
int ensure_vol_lo(thx_ctx* c, const std::vector<char>& used);
void invalidate_vol_lo(thx_ctx* c, int cls);
// The 16-bit local scorers gather from a bricked BF16 copy of each class
// volume.  Element (x, y, z) holds the BF16 complex voxels x and x + 1 of row
// (y, z) and, with kVbRows = 2, also of row ((y + 1) mod V, z): 8 or 16
// bytes, so a trilinear corner set is four or two loads.  Elements are
// grouped in bricks of kVbX x kVbY x kVbZ (one 128-byte cache line), bricks
// x fastest, then y, then z.  The x + 1 voxel of the last column is 0.
#ifndef THX_VB_ROWS
#define THX_VB_ROWS 1
#define THX_VB_X 1
#define THX_VB_Y 4
#define THX_VB_Z 4
#endif
constexpr int kVbRows = THX_VB_ROWS, kVbX = THX_VB_X, kVbY = THX_VB_Y, kVbZ = THX_VB_Z;
constexpr int kVbPer = 16 / kVbRows;   // elements per brick
static_assert(kVbX * kVbY * kVbZ == kVbPer, "a brick is one 128-byte line");
__host__ __device__ inline int vb_nbx(int V) { return (V / 2 + 1 + kVbX - 1) / kVbX; }
__host__ __device__ inline int vb_nby(int V) { return (V + kVbY - 1) / kVbY; }
__host__ __device__ inline int vb_nbz(int V) { return (V + kVbZ - 1) / kVbZ; }
__host__ __device__ inline size_t vb_bytes(int V) {
    return (size_t)vb_nbx(V) * vb_nby(V) * vb_nbz(V) * 128;
}
int launch_vol_bricked(hipStream_t st, const float2* vol, int V, uint32_t* vb);
// 30-bit Morton code of (x, y, z) of a unit quaternion taken with w >= 0
// (q = -q), 10 bits per axis over [-1, 1]: nearby rotations get nearby codes
// except across the w = 0 seam
__device__ inline uint32_t quat_morton_dev(const double* q) {
    const double s = q[0] < 0 ? -1.0 : 1.0;
    uint32_t m = 0, c[3];
    for (int a = 0; a < 3; a++) {
        const double v = fmin(1.0, fmax(-1.0, s * q[a + 1]));
        c[a] = (uint32_t)fmin(1023.0, floor((v + 1.0) * 512.0));
    }
    for (int b = 9; b >= 0; b--)
        for (int a = 0; a < 3; a++) m = (m << 1) | ((c[a] >> b) & 1u);
    return m;
}
// Whole-grid scan (class cls, precision) of the particles idx: their marginals
// land in ctx->lpRes, rows in the order of idx (pinnedIdx: a pinned copy)
int scan_subset(thx_ctx* c, int cls, int precision, const std::vector<int>& idx,
                const int* pinnedIdx = nullptr);
// frees the projectee FFT plans and scratch
void projectee_release(thx_ctx* c);
void images_release(thx_ctx* c);
int alloc_particles(thx_ctx* c, int P);
int choose_local_mode(thx_ctx* c, int precision, int TPAD, int nblk, const LocalBatch& b);
int launch_merge_local(thx_ctx* c, int P, int wtStride, const LocalBatch& b, double* wC, double* wR,
                       double* wT, double* base, int* bestR, int* bestT,
                       int rb = kLocalRotBlock);

// Per-phase scalars of the local search that a replayed graph reads from
// device memory.  With `dev` set, PhaseArg is mul * (*dev) + add and
// PhaseScale is `first` at phase 0 and `rest` after (the phase counter the
// phase's last kernel advances); otherwise both are their `host` value.
struct PhaseArg {
    const int* dev = nullptr;
    int mul = 0, add = 0;
    uint32_t host = 0;
    __device__ uint32_t get() const { return dev ? (uint32_t)(mul * *dev + add) : host; }
};
struct PhaseScale {
    const int* dev = nullptr;
    double first = 0.0, rest = 0.0, host = 0.0;
    __device__ double get() const { return dev ? (*dev == 0 ? first : rest) : host; }
};
inline PhaseArg phase_host(uint32_t v) { PhaseArg a; a.host = v; return a; }
inline PhaseArg phase_dev(const int* d, int mul, int add) {
    PhaseArg a; a.dev = d; a.mul = mul; a.add = add; return a;
}
inline PhaseScale scale_host(double v) { PhaseScale a; a.host = v; return a; }
inline PhaseScale scale_dev(const int* d, double first, double rest) {
    PhaseScale a; a.dev = d; a.first = first; a.rest = rest; return a;
}

// Particle-filter launchers (thx_pf.hip).
int pf_perturb_rot(hipStream_t st, double* quat, const int* off, const int* cnt, int maxCnt,
                   const double* mean, const double* kk, double kmax, double pf, int P,
                   uint64_t seed, uint32_t phase, const double* dGiven);
int pf_perturb_rot(hipStream_t st, double* quat, const int* off, const int* cnt, int maxCnt,
                   const double* mean, const double* kk, double kmax, PhaseScale pf, int P,
                   uint64_t seed, PhaseArg phase);
int pf_perturb_trans(hipStream_t st, double* trans, const int* off, const int* cnt, int maxCnt,
                     const double* s01rho, double pf, int P, uint64_t seed, uint32_t phase);
int pf_sample_mixture(hipStream_t st, const double* mix, int C, int P, int nRot, int nTr,
                      uint64_t seed, uint32_t phase, double* quat, double* qdens,
                      double* trans);
int pf_resample(hipStream_t st, const double* w, const double* u, const int* offIn,
                const int* nIn, const int* offOut, const int* nOut, int P, uint64_t seed,
                uint32_t phase, const double* u0Given, int* idx, double* wNew, double* ess,
                int wShared = 0);
int pf_resample(hipStream_t st, const double* w, const double* u, const int* offIn,
                const int* nIn, const int* offOut, const int* nOut, int P, uint64_t seed,
                PhaseArg phase, int* idx, double* wNew);
int pf_rot_mean(hipStream_t st, const double* quat, const double* w, const int* off,
                const int* cnt, int P, double* mean, double* kk);
constexpr int kResampleMaxN = 8192;
int pf_select_grid(hipStream_t st, const double* mix, int C, int P, const double* gridQ, int R,
                   const double* gridT, int T, int nSelR, int nSelT, int* ridx, float* rlogq,
                   int* tidx, float* spread,
                   float4* gridQf);
// Order of the particles for the selected-pose scorer: by a Morton code of
// the dominant component's mean of each particle's proposal.  tmp == nullptr
// returns the scratch size in *tmpBytes; keyA/keyB/val hold P entries.
// The learned step's per-particle results packed as [score P][ess P] doubles,
// [bestR P][bestT P] ints, [spread P][tail P] floats into dst
int launch_pack_learned(hipStream_t st, const double* score, const double* ess, const int* bestR,
                        const int* bestT, const float* spread, const float* tail, int P,
                        void* dst);
// dst row i = src row idx[i] (rows of rowLen doubles), i < n
int launch_gather_rows(hipStream_t st, const double* src, int rowLen, const int* idx, int n,
                       double* dst);
int pf_select_order(hipStream_t st, const double* mix, int C, int P, uint32_t* keyA,
                    uint32_t* keyB, int* val, int* order, void* tmp, size_t* tmpBytes);
int launch_score_sel(thx_ctx* c, int TPAD, int RB, int nblk, const int* ridx, const int* tidx,
                     int nSelR, int nSelT, float pRu, float pTu, const int* pmap,
                     const int* order, const LocalBatch& b);
int launch_sel_stats(hipStream_t st, const double* wR, const float* rlogq, const int* ridx,
                     const int* tidx, int* bestR, int* bestT, int nSelR, int nSelT, int P,
                     double* ess, float* tail);
int launch_gather_particles(thx_ctx* c, const int* idx, int n, float2* dat2, float* ctf2,
                            float* sig2);

// Philox4x32-10 (Salmon et al., SC'11), shared by host and device code.
struct Philox {
    __host__ __device__ static inline void mulhilo(uint32_t a, uint32_t b,
                                                   uint32_t& hi, uint32_t& lo) {
        uint64_t p = (uint64_t)a * (uint64_t)b;
        hi = (uint32_t)(p >> 32);
        lo = (uint32_t)p;
    }
    __host__ __device__ static inline void round10(uint32_t c[4], uint32_t k0,
                                                   uint32_t k1) {
        for (int r = 0; r < 10; r++) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c[0], hi0, lo0);
            mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
            uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
            c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    // Four 32-bit words for (seed, a, b, stream).
    __host__ __device__ static inline void draw(uint64_t seed, uint32_t a,
                                                uint32_t b, uint32_t stream,
                                                uint32_t out[4]) {
        out[0] = a; out[1] = b; out[2] = stream; out[3] = 0x7468756eu;
        round10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
    }
    // Uniform double in (0, 1) from two words.
    __host__ __device__ static inline double u01(uint32_t hi, uint32_t lo) {
        uint64_t x = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
        x &= ((1ull << 53) - 1);
        return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
    }
};

enum Stream : uint32_t {
    kStreamBlob = 1,
    kStreamPose = 2,
    kStreamShift = 3,
    kStreamCtf = 4,
    kStreamNoise = 5,
    kStreamClass = 6
};

// Hamilton product (quaternion_mul, src/Geometry/Euler.cpp:13-26).
__device__ inline void qmul(const double a[4], const double b[4], double o[4]) {
    const double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w; o[1] = x; o[2] = y; o[3] = z;
}

__device__ inline void normals4(uint64_t seed, uint32_t a, uint32_t b, uint32_t stream,
                                double g[4]) {
    uint32_t w[4], v[4];
    Philox::draw(seed, a, 2 * b, stream, w);
    Philox::draw(seed, a, 2 * b + 1, stream, v);
    const double u1 = Philox::u01(w[0], w[1]), u2 = Philox::u01(w[2], w[3]);
    const double u3 = Philox::u01(v[0], v[1]), u4 = Philox::u01(v[2], v[3]);
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    double s1, c1, s2, c2;
    sincospi(2.0 * u2, &s1, &c1);
    sincospi(2.0 * u4, &s2, &c2);
    g[0] = r1 * c1; g[1] = r1 * s1; g[2] = r2 * c2; g[3] = r2 * s2;
}

// ACG(diag(1, k1, k2, k3)) sample (sampleACG, src/Geometry/DirectionalStat.cpp:
// 33-85): v ~ N(0, I), v = L v with L = chol(diag), normalise.
__device__ inline void acg_identity(const double g[4], double k1, double k2, double k3,
                                    double d[4]) {
    d[0] = g[0];
    d[1] = sqrt(k1) * g[1];
    d[2] = sqrt(k2) * g[2];
    d[3] = sqrt(k3) * g[3];
    const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
    for (int i = 0; i < 4; i++) d[i] /= n;
}


}  // namespace thx
