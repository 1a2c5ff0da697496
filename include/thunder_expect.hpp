/*
 * thunder_expect.hpp -- C++ call surface of THUNDER's expectation step over
 * the C ABI of thunder_expect.h.
 *
 * The reference drives its expectation through three C++ surfaces; this
 * header keeps each of them so that the calling code changes only in the
 * types it passes:
 *   - Projector (include/Projector.h:242-371): setProjectee and
 *     project(Complex*, const dmat33&, iCol, iRow, nPxl, nThread), 3-D;
 *   - Optimiser::expectation() (src/Optimiser.cpp:1438-1681, with
 *     doGlobalSearch :633-1137): the pixel set, the packed particle arrays
 *     (_iCol / _iRow / _nPxl / _datP / _ctfP / _sigRcpP) and the particle
 *     states (_par) in, the states after the local search out;
 *   - the per-particle local-search entry points ExpectLocalRTD,
 *     ExpectLocalPreI3D and ExpectLocalM (gpu/interface/Interface.h:80-87,
 *     107-123, 125-139) with their ManagedCalPoint handle
 *     (gpu/src/ManagedCalPoint.cu:6-104).
 * Everything runs in the library (libthunder_expect.so, which also holds the
 * compiled adapter, csrc/thx_adapter.cpp).  Where the reference prints and
 * calls exit(1), the adapter throws thx::Error with the library's message.
 *
 * Types: Complex is THUNDER's {RFLOAT dat[2]} with RFLOAT = float
 * (include/Complex.h, include/Precision.h:65); a dmat33 is passed as its nine
 * doubles in Eigen's column-major storage (mat.data()).
 */
#ifndef THUNDER_EXPECT_HPP
#define THUNDER_EXPECT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "thunder_expect.h"

namespace thx {

struct Complex {
    float dat[2];
};

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status(status) {}
    int status;
};

/* One GPU and the data resident on it: the pixel set and the packed particle
 * arrays of the expectation, and the class volumes (the per-device state that
 * Optimiser::expectationG sets up, src/Optimiser.cpp:1696-1711, 2197-2342). */
class Device {
public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    /* Optimiser::allocPreCalIdx (src/Optimiser.cpp:7942-7992): the pixel set,
     * the image size N (_para.size) and the padding factor. */
    void setPixels(const int* iCol, const int* iRow, const int* iSig, int nPxl, int size, int pf);
    /* Optimiser::allocPreCal (src/Optimiser.cpp:7994-8121), image-major:
     * datP[nImg * nPxl], ctfP and sigRcpP likewise. */
    void setParticles(const Complex* datP, const float* ctfP, const float* sigRcpP, int nImg);

    int nPxl() const { return nPxl_; }
    int nImg() const { return nImg_; }
    int size() const { return size_; }
    int pf() const { return pf_; }
    thx_ctx* ctx() const { return ctx_; }

private:
    thx_ctx* ctx_ = nullptr;
    int nPxl_ = 0, nImg_ = 0, size_ = 0, pf_ = 2;
};

/* The 3-D Projector (include/Projector.h:242-371) of class cls, its projectee
 * held in cls's slot of the device. */
class Projector {
public:
    Projector(Device& dev, int cls);

    bool isEmpty3D() const { return vdim_ == 0; }
    int pf() const { return pf_; }
    void setPf(int pf) { pf_ = pf; }
    /* LINEAR_INTERP (1) only: the interpolation of the 3-D search
     * (include/Model.h:90-93). */
    int interp() const { return interp_; }
    void setInterp(int interp);
    int maxRadius() const { return maxRadius_; }
    void setMaxRadius(int maxRadius) { maxRadius_ = maxRadius; }

    /* setProjectee(Volume src, nThread) (src/Projector.cpp:123-148): src is the
     * N^3 map's half spectrum, (N/2+1)*N*N, x fastest; the padded, grid-
     * corrected projectee is built on the GPU; maxRadius becomes N/2 - 1. */
    void setProjectee(const Complex* volFT, int N);
    /* The padded projectee itself, as projectee3D() holds it: (V/2+1)*V*V. */
    void setProjectee3D(const Complex* vol, int vdim);
    /* A host copy of projectee3D(). */
    std::vector<Complex> projectee3D() const;

    /* project(Complex* dst, const dmat33& mat, iCol, iRow, nPxl, nThread)
     * (src/Projector.cpp:356-374): dst[i] = the projectee at
     * mat (iCol[i] pf, iRow[i] pf, 0); mat column-major. */
    void project(Complex* dst, const double* mat, const int* iCol, const int* iRow,
                 int nPxl) const;
    /* nR rotations in one launch: dst[r * nPxl + i], mats[9 * r ...]. */
    void project(Complex* dst, const double* mats, int nR, const int* iCol, const int* iRow,
                 int nPxl) const;

private:
    Device* dev_;
    int cls_, pf_ = 2, interp_ = 1, maxRadius_ = 0, vdim_ = 0;
};

/* The members of OptimiserPara (include/Optimiser.h:77-438) the expectation
 * reads, with the 3-D particle filter's defaults. */
struct ExpectPara {
    int k = 1;                          /* classes                                     */
    int mLR = 300, mLT = 25;            /* samples per particle in the local search     */
    double perturbFactorL = 2.0;        /* perturbation of the first local phase        */
    double perturbFactorSGlobal = 0.5;  /* later phases after a global search           */
    double perturbFactorSLocal = 0.5;   /* later phases of a local search               */
    double transS = 1.0;                /* translation search deviation (pixels)        */
    double transSearchFactor = 1.0;     /* scanning minimum of the shift deviation      */
    int mS = 5000;                      /* global rotation sampling: scanMinStdR       */
    bool globalSearch = true;           /* _searchType == SEARCH_TYPE_GLOBAL            */
    int precision = THX_FP32;           /* arithmetic of the cross term                 */
    int maxPhase = 100;                 /* MAX_N_PHASE_PER_ITER                         */
    uint64_t seed = 1;                  /* random streams (Philox4x32-10)              */
};

/* What Particle::load / store exchange (include/Particle.h) for nImg
 * particles, FP64. */
struct ParticleStates {
    int mLR = 0, mLT = 0;
    std::vector<int> cls;            /* [nImg]                 */
    std::vector<double> quat;        /* [nImg * mLR * 4]       */
    std::vector<double> trans;       /* [nImg * mLT * 2] pixels */
    std::vector<double> wR, wT;      /* [nImg * mLR], [nImg * mLT] */
    std::vector<double> k;           /* [nImg * 3]             */
    std::vector<double> s01;         /* [nImg * 2]             */
    std::vector<double> peakR;       /* [nImg]                 */
    std::vector<double> topR, topT;  /* [nImg * 4], [nImg * 2] */
    std::vector<int> nPhase;         /* [nImg], -1: not converged */
};

/* Optimiser::expectation() on the device's pixel set and particles. */
class Expectation {
public:
    explicit Expectation(Device& dev) : dev_(&dev) {}
    /* The global search's pose grid (Particle::reset, src/Particle.cpp:
     * 87-169): quaternions (w, x, y, z) [nR * 4], shifts [nT * 2] in pixels,
     * prior weights (uniform when null). */
    void setGlobalGrid(const double* quat, int nR, const double* trans, int nT,
                       const double* pR = nullptr, const double* pT = nullptr);
    /* With para.globalSearch, the global scan of classes 0..k-1 and the
     * post-scan update (src/Optimiser.cpp:633-1137) make the states; otherwise
     * par holds the states to start from (cls, quat, trans, k, s01, peakR).
     * The local phases then run to convergence (:1438-1681).  par receives
     * the final states; returns the number of phases run. */
    int expectation(const ExpectPara& para, ParticleStates& par);

private:
    Device* dev_;
    int nR_ = 0, nT_ = 0;
};

/* The per-particle scratch of the reference's local search
 * (ManagedCalPoint, gpu/src/ManagedCalPoint.cu:6-104): the particle's
 * rotations, shifts and their prior weights between the three calls below. */
class ManagedCalPoint {
public:
    /* mode 1 (3-D); cSearch 0 (no CTF search: the 2-D and CTF-search
     * branches are not built); mD must be 1. */
    void Init(int mode, int cSearch, Device& dev, int nR, int nT, int mD, int npxl);
    int getNR() const { return nR_; }
    int getNT() const { return nT_; }
    int getMD() const { return 1; }
    int getCSearch() const { return cSearch_; }

private:
    friend void ExpectLocalRTD(Device&, ManagedCalPoint*, const double*, const double*,
                               const double*, const double*, const double*, const double*);
    friend void ExpectLocalPreI3D(Device&, int, int, ManagedCalPoint*, const float*,
                                  const float*, const int*, const int*, float, float, float,
                                  float, int, int, int, int, int);
    friend void ExpectLocalM(Device&, int, ManagedCalPoint*, float*, float*, float*, float*,
                             double, int);
    int nR_ = 0, nT_ = 0, npxl_ = 0, cSearch_ = 0, cls_ = 0;
    bool ready_ = false;
    std::vector<double> rot_, trans_, oldR_, oldT_;
    double oldD_ = 1.0;
};

/* ExpectLocalRTD (gpu/interface/Interface.h:80-87): the particle's prior
 * weights oldR[nR], oldT[nT], oldD[1], its shifts trans[nT * 2] and
 * quaternions rot[nR * 4]; dpara (CTF search) unused. */
void ExpectLocalRTD(Device& dev, ManagedCalPoint* mcp, const double* oldR, const double* oldT,
                    const double* oldD, const double* trans, const double* rot,
                    const double* dpara);
/* ExpectLocalPreI3D (Interface.h:107-123): projections against class cls
 * (the reference's ManagedArrayTexture) for particle datShift.  Slices are
 * projected inside the scorer, so this call records the class and checks
 * the sizes; the CTF-search inputs (devdefO, devfreQ, phaseShift, conT, k1,
 * k2) are unused and deviCol / deviRow are the device's pixel set. */
void ExpectLocalPreI3D(Device& dev, int datShift, int cls, ManagedCalPoint* mcp,
                       const float* devdefO, const float* devfreQ, const int* deviCol,
                       const int* deviRow, float phaseShift, float conT, float k1, float k2,
                       int pf, int idim, int vdim, int npxl, int interp);
/* ExpectLocalM (Interface.h:125-139): scores particle datShift against its
 * nR x nT poses and writes, with w = exp(dvp - max dvp) (kernel_UpdateWL,
 * gpu/src/Kernel.cu:1230-1345):
 *   wC = sum w oldR oldT oldD,  wR[r] = sum_t w oldT oldD oldC,
 *   wT[t] = sum_r w oldR oldD oldC,  wD[0] = sum w oldR oldT oldC. */
void ExpectLocalM(Device& dev, int datShift, ManagedCalPoint* mcp, float* wC, float* wR,
                  float* wT, float* wD, double oldC, int npxl);

}  // namespace thx

#endif
