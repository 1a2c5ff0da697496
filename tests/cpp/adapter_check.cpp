// Exercises the C++ adapter (include/thunder_expect.hpp) through
// libthunder_expect.so on inputs written by tests/test_gpu_adapter.py, and
// writes what it computes for that test to compare with the oracle and with
// the C ABI called directly.
//
//   adapter_check DIR
//
// DIR/meta.txt holds "name value" lines; arrays are raw little-endian files
// named after the array.  Outputs go to DIR/out_*.bin.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "thunder_expect.hpp"

static std::string g_dir;

template <class T>
static std::vector<T> load(const std::string& name, size_t n) {
    std::ifstream f(g_dir + "/" + name, std::ios::binary);
    if (!f) throw std::runtime_error("missing input " + name);
    std::vector<T> v(n);
    f.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n);
    if ((size_t)f.gcount() != sizeof(T) * n) throw std::runtime_error("short input " + name);
    return v;
}

template <class T>
static void save(const std::string& name, const T* p, size_t n) {
    std::ofstream f(g_dir + "/out_" + name + ".bin", std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), sizeof(T) * n);
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: adapter_check DIR\n");
        return 2;
    }
    g_dir = argv[1];
    std::map<std::string, long long> m;
    {
        std::ifstream f(g_dir + "/meta.txt");
        std::string k;
        long long v;
        while (f >> k >> v) m[k] = v;
    }
    try {
        const int N = (int)m.at("N"), pf = (int)m.at("pf"), V = N * pf, K = (int)m.at("K");
        const int P = (int)m.at("P"), R = (int)m.at("R"), T = (int)m.at("T");
        const int nM = (int)m.at("nM"), Ns = (int)m.at("Ns"), nL = (int)m.at("nL");
        thx::Device dev(0);
        const auto iCol = load<int>("iCol", K), iRow = load<int>("iRow", K);
        const auto iSig = load<int>("iSig", K);
        dev.setPixels(iCol.data(), iRow.data(), iSig.data(), K, N, pf);
        // Projector: a padded projectee, batched and single-rotation projection
        const size_t nv = (size_t)(V / 2 + 1) * V * V;
        const auto vol = load<thx::Complex>("vol", nv);
        thx::Projector prj(dev, 0);
        prj.setProjectee3D(vol.data(), V);
        const auto mats = load<double>("mats", 9 * (size_t)nM);
        std::vector<thx::Complex> out((size_t)nM * K), one(K);
        prj.project(out.data(), mats.data(), nM, iCol.data(), iRow.data(), K);
        prj.project(one.data(), mats.data(), iCol.data(), iRow.data(), K);
        save("project", out.data(), out.size());
        save("project1", one.data(), one.size());
        // Projector::setProjectee from a small map's spectrum (class slot 1)
        const auto small = load<thx::Complex>("volft", (size_t)(Ns / 2 + 1) * Ns * Ns);
        thx::Projector prj1(dev, 1);
        prj1.setProjectee(small.data(), Ns);
        const auto padded = prj1.projectee3D();
        save("projectee", padded.data(), padded.size());
        int mr = prj1.maxRadius();
        save("maxradius", &mr, 1);
        // particles
        const auto dat = load<thx::Complex>("dat", (size_t)P * K);
        const auto ctf = load<float>("ctf", (size_t)P * K), sig = load<float>("sig", (size_t)P * K);
        dev.setParticles(dat.data(), ctf.data(), sig.data(), P);
        // per-particle local search shims
        const auto lp = load<int>("local_idx", 4 * (size_t)nL);   // particle, nR, nT, offsets
        const auto lrot = load<double>("local_rot", (size_t)m.at("local_nrot") * 4);
        const auto ltr = load<double>("local_trans", (size_t)m.at("local_ntrans") * 2);
        const auto loR = load<double>("local_oldR", (size_t)m.at("local_nrot"));
        const auto loT = load<double>("local_oldT", (size_t)m.at("local_ntrans"));
        const auto loC = load<double>("local_oldC", nL);
        size_t ro = 0, to = 0;
        for (int i = 0; i < nL; i++) {
            const int l = lp[4 * i], nR = lp[4 * i + 1], nT = lp[4 * i + 2];
            thx::ManagedCalPoint mcp;
            mcp.Init(1, 0, dev, nR, nT, 1, K);
            const double oldD = 1.0;
            thx::ExpectLocalRTD(dev, &mcp, loR.data() + ro, loT.data() + to, &oldD,
                                ltr.data() + 2 * to, lrot.data() + 4 * ro, nullptr);
            thx::ExpectLocalPreI3D(dev, l, 0, &mcp, nullptr, nullptr, nullptr, nullptr, 0.f, 0.f,
                                   0.f, 0.f, pf, N, V, K, 1);
            std::vector<float> w(2 + nR + nT);
            thx::ExpectLocalM(dev, l, &mcp, &w[0], &w[2], &w[2 + nR], &w[1], loC[i], K);
            save("local_" + std::to_string(i), w.data(), w.size());
            ro += nR;
            to += nT;
        }
        // Optimiser::expectation: global scan + post-scan + local phases
        const auto gq = load<double>("grid_quat", 4 * (size_t)R);
        const auto gt = load<double>("grid_trans", 2 * (size_t)T);
        thx::Expectation ex(dev);
        ex.setGlobalGrid(gq.data(), R, gt.data(), T);
        thx::ExpectPara para;
        para.k = 1;
        para.mLR = (int)m.at("mLR");
        para.mLT = (int)m.at("mLT");
        para.mS = R;
        para.perturbFactorL = 2.0;
        para.perturbFactorSGlobal = 0.5;
        para.transS = 1.0;
        para.maxPhase = (int)m.at("maxPhase");
        para.precision = (int)m.at("precision");
        para.seed = (uint64_t)m.at("seed");
        thx::ParticleStates st;
        int phases = ex.expectation(para, st);
        save("phases", &phases, 1);
        save("quat", st.quat.data(), st.quat.size());
        save("trans", st.trans.data(), st.trans.size());
        save("k", st.k.data(), st.k.size());
        save("s01", st.s01.data(), st.s01.size());
        save("topR", st.topR.data(), st.topR.size());
        save("topT", st.topT.data(), st.topT.size());
        save("nPhase", st.nPhase.data(), st.nPhase.size());
        // a state error is an exception with the library's message
        try {
            thx::Projector bad(dev, 5);
            bad.project(one.data(), mats.data(), iCol.data(), iRow.data(), K);
            std::printf("missing error\n");
            return 1;
        } catch (const thx::Error& e) {
            std::printf("error as expected: %s\n", e.what());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter_check: %s\n", e.what());
        return 1;
    }
    std::printf("adapter_check done\n");
    return 0;
}
