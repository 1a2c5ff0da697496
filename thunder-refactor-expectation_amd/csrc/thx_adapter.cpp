// C++ call surface of the reference's expectation step (Projector,
// Optimiser::expectation, the per-particle ExpectLocal* functions), built on
// the C ABI of thunder_expect.h.  Host code only: every computation is one of
// the library's entry points.

#include "thunder_expect.hpp"

#include <cmath>
#include <cstring>

namespace thx {

namespace {

void check(int status, const char* what) {
    if (status != THX_OK) throw Error(status, std::string(what) + ": " + thx_last_error());
}

void require(bool cond, const char* msg) {
    if (!cond) throw Error(THX_ERR_ARG, msg);
}

}  // namespace

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------

Device::Device(int device) { check(thx_ctx_create(device, &ctx_), "thx_ctx_create"); }

Device::~Device() {
    if (ctx_) thx_ctx_destroy(ctx_);
}

void Device::setPixels(const int* iCol, const int* iRow, const int* iSig, int nPxl, int size,
                       int pf) {
    check(thx_ctx_set_pixels(ctx_, iCol, iRow, iSig, nPxl, size, pf), "setPixels");
    nPxl_ = nPxl;
    size_ = size;
    pf_ = pf;
    nImg_ = 0;   // the particle arrays depend on the pixel set
}

void Device::setParticles(const Complex* datP, const float* ctfP, const float* sigRcpP,
                          int nImg) {
    check(thx_ctx_set_particles(ctx_, reinterpret_cast<const float*>(datP), ctfP, sigRcpP, nImg),
          "setParticles");
    nImg_ = nImg;
}

// ---------------------------------------------------------------------------
// Projector
// ---------------------------------------------------------------------------

Projector::Projector(Device& dev, int cls) : dev_(&dev), cls_(cls) {
    require(cls >= 0, "Projector: class index must be >= 0");
}

void Projector::setInterp(int interp) {
    require(interp == 1, "Projector: only LINEAR_INTERP (1) is supported in 3-D");
    interp_ = interp;
}

void Projector::setProjectee(const Complex* volFT, int N) {
    check(thx_ctx_set_projectee(dev_->ctx(), cls_, reinterpret_cast<const float*>(volFT), N, pf_,
                                nullptr),
          "Projector::setProjectee");
    vdim_ = pf_ * N;
    maxRadius_ = N / 2 - 1;
}

void Projector::setProjectee3D(const Complex* vol, int vdim) {
    check(thx_ctx_set_volume(dev_->ctx(), cls_, reinterpret_cast<const float*>(vol), vdim),
          "Projector::setProjectee3D");
    vdim_ = vdim;
    maxRadius_ = vdim / pf_ / 2 - 1;
}

std::vector<Complex> Projector::projectee3D() const {
    require(vdim_ > 0, "Projector::projectee3D: no projectee set");
    std::vector<Complex> out((size_t)(vdim_ / 2 + 1) * vdim_ * vdim_);
    check(thx_ctx_get_volume(dev_->ctx(), cls_, reinterpret_cast<float*>(out.data())),
          "Projector::projectee3D");
    return out;
}

void Projector::project(Complex* dst, const double* mat, const int* iCol, const int* iRow,
                        int nPxl) const {
    project(dst, mat, 1, iCol, iRow, nPxl);
}

void Projector::project(Complex* dst, const double* mats, int nR, const int* iCol,
                        const int* iRow, int nPxl) const {
    require(vdim_ > 0, "Projector::project: no projectee set");
    check(thx_ctx_project(dev_->ctx(), cls_, pf_, mats, nR, iCol, iRow, nPxl,
                          reinterpret_cast<float*>(dst)),
          "Projector::project");
}

// ---------------------------------------------------------------------------
// Expectation
// ---------------------------------------------------------------------------

void Expectation::setGlobalGrid(const double* quat, int nR, const double* trans, int nT,
                                const double* pR, const double* pT) {
    require(quat && trans && nR > 0 && nT > 0, "setGlobalGrid: bad argument");
    std::vector<double> uR, uT;
    if (!pR) {
        uR.assign(nR, 1.0 / nR);
        pR = uR.data();
    }
    if (!pT) {
        uT.assign(nT, 1.0 / nT);
        pT = uT.data();
    }
    check(thx_ctx_set_grid(dev_->ctx(), quat, nR, trans, nT, pR, pT), "setGlobalGrid");
    nR_ = nR;
    nT_ = nT;
}

int Expectation::expectation(const ExpectPara& para, ParticleStates& par) {
    thx_ctx* c = dev_->ctx();
    const int P = dev_->nImg();
    require(P > 0 && dev_->nPxl() > 0, "expectation: pixels and particles must be set");
    require(para.k >= 1 && para.mLR >= 2 && para.mLT >= 2, "expectation: bad parameters");
    const std::vector<double> transS(P, para.transS);
    const double pfS = para.globalSearch ? para.perturbFactorSGlobal : para.perturbFactorSLocal;
    if (para.globalSearch) {
        require(nR_ > 0, "expectation: no global grid set");
        for (int k = 0; k < para.k; k++)
            check(thx_ctx_score_global(c, k, para.precision), "expectation: global scan");
        // scanning minima (src/Optimiser.cpp:668-686): mS^(-1/3) for the 3-D
        // rotations, 1 / chi2_2 Qinv(INIT_OUTSIDE_CONFIDENCE_AREA = 0.5) /
        // sqrt(transSearchFactor pi) for the shifts, chi2_2 Qinv(q) = -2 ln q
        const double minStdR = std::pow((double)para.mS, -1.0 / 3);
        const double minStdT =
            1.0 / (-2.0 * std::log(0.5)) / std::sqrt(para.transSearchFactor * M_PI);
        par.cls.assign(P, 0);
        if (para.k > 1)
            check(thx_pf_from_global_classes(c, para.k, para.mLR, para.mLT, minStdR, minStdT, pfS,
                                             transS.data(), para.seed, par.cls.data()),
                  "expectation: post-scan update");
        else
            check(thx_pf_from_global(c, 0, para.mLR, para.mLT, minStdR, minStdT, pfS,
                                     transS.data(), para.seed),
                  "expectation: post-scan update");
    } else {
        require(par.mLR == para.mLR && par.mLT == para.mLT &&
                    par.quat.size() == (size_t)P * para.mLR * 4 &&
                    par.trans.size() == (size_t)P * para.mLT * 2 && par.k.size() == (size_t)P * 3 &&
                    par.s01.size() == (size_t)P * 2 && par.peakR.size() == (size_t)P,
                "expectation: particle states do not match the particles and para");
        if (par.cls.size() != (size_t)P) par.cls.assign(P, 0);
        check(thx_pf_set_state(c, para.mLR, para.mLT, par.quat.data(), par.trans.data(),
                               par.k.data(), par.s01.data(), par.peakR.data(), transS.data()),
              "expectation: particle states");
    }
    // the local phases take FP32, BF16 or FP16; an FP8 global scan is
    // followed by FP16 local phases (per-shell scales like FP8's)
    const int localPrec = para.precision == THX_FP8 ? THX_FP16 : para.precision;
    int phases = 0;
    check(thx_pf_refine(c, par.cls.data(), localPrec, para.globalSearch ? 1 : 0,
                        para.perturbFactorL, pfS, para.maxPhase, para.seed, &phases),
          "expectation: local search");
    par.mLR = para.mLR;
    par.mLT = para.mLT;
    par.quat.resize((size_t)P * para.mLR * 4);
    par.trans.resize((size_t)P * para.mLT * 2);
    par.wR.resize((size_t)P * para.mLR);
    par.wT.resize((size_t)P * para.mLT);
    par.k.resize((size_t)P * 3);
    par.s01.resize((size_t)P * 2);
    par.topR.resize((size_t)P * 4);
    par.topT.resize((size_t)P * 2);
    par.peakR.resize(P);
    par.nPhase.resize(P);
    check(thx_pf_get_state(c, par.quat.data(), par.trans.data(), par.wR.data(), par.wT.data(),
                           par.k.data(), par.s01.data(), par.topR.data(), par.topT.data(),
                           par.peakR.data(), par.nPhase.data()),
          "expectation: particle states");
    return phases;
}

// ---------------------------------------------------------------------------
// Per-particle local search (ExpectLocalRTD / PreI3D / M)
// ---------------------------------------------------------------------------

void ManagedCalPoint::Init(int mode, int cSearch, Device& dev, int nR, int nT, int mD,
                           int npxl) {
    (void)dev;
    require(mode == 1, "ManagedCalPoint::Init: only the 3-D mode (1) is built");
    require(cSearch != 2 && mD == 1,
            "ManagedCalPoint::Init: the CTF-search branch (cSearch 2, mD > 1) is not built");
    require(nR >= 1 && nR <= 8192 && nT >= 1 && nT <= 128 && npxl >= 1,
            "ManagedCalPoint::Init: nR must be in [1, 8192] and nT in [1, 128]");
    nR_ = nR;
    nT_ = nT;
    npxl_ = npxl;
    cSearch_ = cSearch;
    rot_.assign(4 * (size_t)nR, 0.0);
    trans_.assign(2 * (size_t)nT, 0.0);
    oldR_.assign(nR, 0.0);
    oldT_.assign(nT, 0.0);
    ready_ = false;
}

void ExpectLocalRTD(Device& dev, ManagedCalPoint* mcp, const double* oldR, const double* oldT,
                    const double* oldD, const double* trans, const double* rot,
                    const double* dpara) {
    (void)dev;
    (void)dpara;
    require(mcp && mcp->nR_ > 0 && oldR && oldT && trans && rot,
            "ExpectLocalRTD: bad argument (ManagedCalPoint::Init first)");
    std::memcpy(mcp->oldR_.data(), oldR, sizeof(double) * mcp->nR_);
    std::memcpy(mcp->oldT_.data(), oldT, sizeof(double) * mcp->nT_);
    std::memcpy(mcp->rot_.data(), rot, sizeof(double) * 4 * mcp->nR_);
    std::memcpy(mcp->trans_.data(), trans, sizeof(double) * 2 * mcp->nT_);
    mcp->oldD_ = oldD ? oldD[0] : 1.0;
    mcp->ready_ = false;
}

void ExpectLocalPreI3D(Device& dev, int datShift, int cls, ManagedCalPoint* mcp,
                       const float* devdefO, const float* devfreQ, const int* deviCol,
                       const int* deviRow, float phaseShift, float conT, float k1, float k2,
                       int pf, int idim, int vdim, int npxl, int interp) {
    (void)devdefO;
    (void)devfreQ;
    (void)deviCol;
    (void)deviRow;
    (void)phaseShift;
    (void)conT;
    (void)k1;
    (void)k2;
    require(mcp && mcp->nR_ > 0, "ExpectLocalPreI3D: ManagedCalPoint::Init first");
    require(datShift >= 0 && datShift < dev.nImg(), "ExpectLocalPreI3D: datShift out of range");
    require(npxl == dev.nPxl() && idim == dev.size() && pf == dev.pf() && vdim == pf * idim,
            "ExpectLocalPreI3D: sizes differ from the device's pixel set");
    require(interp == 1, "ExpectLocalPreI3D: only LINEAR_INTERP (1) is supported");
    require(cls >= 0, "ExpectLocalPreI3D: bad class");
    mcp->cls_ = cls;
    mcp->ready_ = true;
}

void ExpectLocalM(Device& dev, int datShift, ManagedCalPoint* mcp, float* wC, float* wR,
                  float* wT, float* wD, double oldC, int npxl) {
    require(mcp && mcp->ready_, "ExpectLocalM: ExpectLocalRTD and ExpectLocalPreI3D first");
    require(wC && wR && wT && wD, "ExpectLocalM: null output");
    require(npxl == dev.nPxl(), "ExpectLocalM: npxl differs from the device's pixel set");
    const int P = dev.nImg();
    require(datShift >= 0 && datShift < P, "ExpectLocalM: datShift out of range");
    // one particle of the device's batch: every other particle is skipped
    std::vector<int> nR(P, 0), nT(P, 0), cls(P, mcp->cls_);
    nR[datShift] = mcp->nR_;
    nT[datShift] = mcp->nT_;
    check(thx_ctx_score_local(dev.ctx(), cls.data(), mcp->rot_.data(), nR.data(),
                              mcp->trans_.data(), nT.data(), mcp->oldR_.data(),
                              mcp->oldT_.data(), THX_FP32),
          "ExpectLocalM");
    std::vector<double> c(P), r(mcp->nR_), t(mcp->nT_);
    check(thx_ctx_get_local(dev.ctx(), c.data(), r.data(), t.data(), nullptr, nullptr, nullptr),
          "ExpectLocalM");
    // c = sum w oldR oldT, r = sum_t w oldT, t = sum_r w oldR (w relative to
    // the particle's largest dvp); the reference's weights carry oldD and oldC
    const double d = mcp->oldD_;
    wC[0] = (float)(c[datShift] * d);
    for (int i = 0; i < mcp->nR_; i++) wR[i] = (float)(r[i] * d * oldC);
    for (int i = 0; i < mcp->nT_; i++) wT[i] = (float)(t[i] * d * oldC);
    wD[0] = (float)(c[datShift] * oldC);
}

}  // namespace thx
