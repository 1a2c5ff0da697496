// On-disk formats of the expectation's inputs and outputs: the .thu particle
// metadata (one line of 27 fields per particle: Database, src/Database.cpp:
// 100-300, fields include/Database.h:12-88; written by
// Optimiser::saveDatabase, src/Optimiser.cpp:8167-8268) and MRC image stacks
// (ImageFile, src/Image/ImageFile.cpp:174-300: a 1024-byte header, nsymbt
// bytes of extended header, then nz slices of nx * ny values of mode 0
// (int8), 1 (int16) or 2 (float32), x fastest).  Host code only.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "thunder_expect.h"

namespace thx {
void set_error(const std::string& msg);
void clear_error();
}  // namespace thx

using thx::clear_error;
using thx::set_error;

namespace {

constexpr int kThuFields = 27;

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// whitespace-separated tokens of one line (strtok with " " in the reference;
// tabs and the newline are separators here too)
std::vector<std::string> tokens(const char* line) {
    std::vector<std::string> out;
    const char* p = line;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
        if (!*p) break;
        const char* q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
        out.emplace_back(p, q);
        p = q;
    }
    return out;
}

bool read_lines(const char* path, std::vector<std::string>& lines) {
    File f(path, "r");
    if (!f.f) return false;
    std::string cur;
    char buf[4096];
    while (std::fgets(buf, sizeof(buf), f.f)) {
        cur += buf;
        if (!cur.empty() && cur.back() == '\n') {
            lines.push_back(cur);
            cur.clear();
        }
    }
    if (!cur.empty()) lines.push_back(cur);
    return true;
}

int parse_record(const std::vector<std::string>& t, thx_thu_record* r) {
    if ((int)t.size() < kThuFields) return -1;
    std::memset(r, 0, sizeof(*r));
    for (int i = 0; i < 7; i++) r->ctf[i] = (float)std::atof(t[i].c_str());
    if (t[7].size() >= sizeof(r->particle) || t[8].size() >= sizeof(r->micrograph)) return -2;
    std::memcpy(r->particle, t[7].c_str(), t[7].size() + 1);
    std::memcpy(r->micrograph, t[8].c_str(), t[8].size() + 1);
    r->coord[0] = std::atof(t[9].c_str());
    r->coord[1] = std::atof(t[10].c_str());
    r->group = std::atoi(t[11].c_str());
    r->cls = std::atoi(t[12].c_str());
    for (int i = 0; i < 4; i++) r->quat[i] = std::atof(t[13 + i].c_str());
    for (int i = 0; i < 3; i++) r->k[i] = std::atof(t[17 + i].c_str());
    for (int i = 0; i < 2; i++) r->trans[i] = std::atof(t[20 + i].c_str());
    for (int i = 0; i < 2; i++) r->stdTrans[i] = std::atof(t[22 + i].c_str());
    r->df = std::atof(t[24].c_str());
    r->stdDf = std::atof(t[25].c_str());
    r->score = std::atof(t[26].c_str());
    return 0;
}

// the 1024-byte MRC header: the words the reader and writer use
struct MrcHeader {
    int32_t nx, ny, nz, mode;
    int32_t nsymbt;
};

int mrc_header(FILE* f, MrcHeader& h, const char* path) {
    unsigned char b[1024];
    if (std::fseek(f, 0, SEEK_SET) != 0 || std::fread(b, 1, 1024, f) != 1024) {
        set_error(std::string("mrc: cannot read the header of ") + path);
        return THX_ERR_ARG;
    }
    int32_t w[256];
    std::memcpy(w, b, 1024);
    h.nx = w[0];
    h.ny = w[1];
    h.nz = w[2];
    h.mode = w[3];
    h.nsymbt = w[23];
    if (h.nx <= 0 || h.ny <= 0 || h.nz <= 0 || h.nsymbt < 0) {
        set_error(std::string("mrc: bad dimensions in ") + path);
        return THX_ERR_ARG;
    }
    if (h.mode != 0 && h.mode != 1 && h.mode != 2) {
        set_error(std::string("mrc: mode must be 0, 1 or 2 (int8, int16, float32) in ") + path);
        return THX_ERR_RANGE;
    }
    return THX_OK;
}

int mrc_slices(FILE* f, const MrcHeader& h, int first, int count, float* out, const char* path) {
    if (first < 0 || count < 1 || first + count > h.nz) {
        set_error(std::string("mrc: slices out of range in ") + path);
        return THX_ERR_RANGE;
    }
    const size_t n = (size_t)h.nx * h.ny * count;
    const int bytes = h.mode == 0 ? 1 : (h.mode == 1 ? 2 : 4);
    const long long off = 1024ll + h.nsymbt + (long long)h.nx * h.ny * bytes * first;
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) {
        set_error(std::string("mrc: seek failed in ") + path);
        return THX_ERR_ARG;
    }
    if (h.mode == 2) {
        if (std::fread(out, 4, n, f) != n) {
            set_error(std::string("mrc: short file ") + path);
            return THX_ERR_ARG;
        }
        return THX_OK;
    }
    std::vector<unsigned char> raw(n * bytes);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) {
        set_error(std::string("mrc: short file ") + path);
        return THX_ERR_ARG;
    }
    for (size_t i = 0; i < n; i++) {
        if (h.mode == 0) {
            out[i] = (float)(int8_t)raw[i];
        } else {
            int16_t v;
            std::memcpy(&v, raw.data() + 2 * i, 2);
            out[i] = (float)v;
        }
    }
    return THX_OK;
}

}  // namespace

extern "C" {

int thx_thu_read(const char* path, thx_thu_record* rec, int cap, int* n) {
    clear_error();
    if (!path || !n || (cap > 0 && !rec)) {
        set_error("thu_read: bad argument");
        return THX_ERR_ARG;
    }
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) {
        set_error(std::string("thu_read: cannot open ") + path);
        return THX_ERR_ARG;
    }
    int count = 0;
    for (size_t i = 0; i < lines.size(); i++) {
        const auto t = tokens(lines[i].c_str());
        if (t.empty()) continue;   // blank line
        if ((int)t.size() < kThuFields) {
            set_error("thu_read: line " + std::to_string(i + 1) + " of " + path + " has " +
                      std::to_string(t.size()) + " fields, 27 expected");
            return THX_ERR_ARG;
        }
        if (count < cap && parse_record(t, rec + count) != 0) {
            set_error("thu_read: a path on line " + std::to_string(i + 1) + " is too long");
            return THX_ERR_RANGE;
        }
        count++;
    }
    *n = count;
    return THX_OK;
}

int thx_thu_write(const char* path, const thx_thu_record* rec, int n, int append) {
    clear_error();
    if (!path || n < 0 || (n > 0 && !rec)) {
        set_error("thu_write: bad argument");
        return THX_ERR_ARG;
    }
    File f(path, append ? "a" : "w");
    if (!f.f) {
        set_error(std::string("thu_write: cannot open ") + path);
        return THX_ERR_ARG;
    }
    for (int l = 0; l < n; l++) {
        const thx_thu_record& r = rec[l];
        // the field order and %18.9lf of Optimiser::saveDatabase
        if (std::fprintf(f.f,
                         "%18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %s %s %18.9lf "
                         "%18.9lf %6d %6d %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf "
                         "%18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf %18.9lf\n",
                         (double)r.ctf[0], (double)r.ctf[1], (double)r.ctf[2], (double)r.ctf[3],
                         (double)r.ctf[4], (double)r.ctf[5], (double)r.ctf[6], r.particle,
                         r.micrograph, r.coord[0], r.coord[1], r.group, r.cls, r.quat[0],
                         r.quat[1], r.quat[2], r.quat[3], r.k[0], r.k[1], r.k[2], r.trans[0],
                         r.trans[1], r.stdTrans[0], r.stdTrans[1], r.df, r.stdDf, r.score) < 0) {
            set_error(std::string("thu_write: write failed on ") + path);
            return THX_ERR_ARG;
        }
    }
    return THX_OK;
}

int thx_mrc_info(const char* path, int* nx, int* ny, int* nz, int* mode) {
    clear_error();
    if (!path) {
        set_error("mrc_info: null path");
        return THX_ERR_ARG;
    }
    File f(path, "rb");
    if (!f.f) {
        set_error(std::string("mrc_info: cannot open ") + path);
        return THX_ERR_ARG;
    }
    MrcHeader h;
    const int s = mrc_header(f.f, h, path);
    if (s) return s;
    if (nx) *nx = h.nx;
    if (ny) *ny = h.ny;
    if (nz) *nz = h.nz;
    if (mode) *mode = h.mode;
    return THX_OK;
}

int thx_mrc_read(const char* path, int first, int count, float* out) {
    clear_error();
    if (!path || !out) {
        set_error("mrc_read: bad argument");
        return THX_ERR_ARG;
    }
    File f(path, "rb");
    if (!f.f) {
        set_error(std::string("mrc_read: cannot open ") + path);
        return THX_ERR_ARG;
    }
    MrcHeader h;
    int s = mrc_header(f.f, h, path);
    if (s) return s;
    return mrc_slices(f.f, h, first, count, out, path);
}

int thx_mrc_write(const char* path, const float* data, int nx, int ny, int nz, float pixelSize) {
    clear_error();
    if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0) {
        set_error("mrc_write: bad argument");
        return THX_ERR_ARG;
    }
    File f(path, "wb");
    if (!f.f) {
        set_error(std::string("mrc_write: cannot open ") + path);
        return THX_ERR_ARG;
    }
    // the header ImageFile::fillMRCHeader writes (mode 2), with the cell in
    // Angstrom when a pixel size is given, the density range and the stamp
    int32_t w[256] = {};
    float fw[256] = {};
    const size_t n = (size_t)nx * ny * nz;
    double mn = data[0], mx = data[0], sum = 0.0, ss = 0.0;
    for (size_t i = 0; i < n; i++) {
        mn = data[i] < mn ? data[i] : mn;
        mx = data[i] > mx ? data[i] : mx;
        sum += data[i];
    }
    const double mean = sum / n;
    for (size_t i = 0; i < n; i++) ss += (data[i] - mean) * (data[i] - mean);
    w[0] = nx; w[1] = ny; w[2] = nz; w[3] = 2;
    w[7] = nx; w[8] = ny; w[9] = nz;                      // mx, my, mz
    const float px = pixelSize > 0.f ? pixelSize : 1.f;
    fw[10] = nx * px; fw[11] = ny * px; fw[12] = nz * px;   // cella
    fw[13] = fw[14] = fw[15] = 90.f;                       // cellb
    w[16] = 1; w[17] = 2; w[18] = 3;                       // mapc, mapr, maps
    fw[19] = (float)mn; fw[20] = (float)mx; fw[21] = (float)mean;
    w[22] = 1;                                             // ispg
    w[23] = 0;                                             // nsymbt
    unsigned char b[1024];
    for (int i = 0; i < 256; i++) {
        const bool isFloat = (i >= 10 && i <= 15) || (i >= 19 && i <= 21);
        if (isFloat) std::memcpy(b + 4 * i, &fw[i], 4);
        else std::memcpy(b + 4 * i, &w[i], 4);
    }
    std::memcpy(b + 208, "MAP ", 4);
    const unsigned char stamp[4] = {0x44, 0x44, 0x00, 0x00};   // little-endian
    std::memcpy(b + 212, stamp, 4);
    const float rms = (float)std::sqrt(ss / n);
    std::memcpy(b + 216, &rms, 4);
    if (std::fwrite(b, 1, 1024, f.f) != 1024 || std::fwrite(data, 4, n, f.f) != n) {
        set_error(std::string("mrc_write: write failed on ") + path);
        return THX_ERR_ARG;
    }
    return THX_OK;
}

int thx_ctx_load_thu(thx_ctx* ctx, const char* thuPath, const char* parPrefix, int first,
                     int count, float pixelSize, float bgRadius, int softMask, float* msOut) {
    clear_error();
    if (!ctx || !thuPath || first < 0 || count < 1 || pixelSize <= 0.f) {
        set_error("load_thu: bad argument");
        return THX_ERR_ARG;
    }
    int n = 0;
    int s = thx_thu_read(thuPath, nullptr, 0, &n);
    if (s) return s;
    if (first + count > n) {
        set_error("load_thu: the file has " + std::to_string(n) + " particles");
        return THX_ERR_RANGE;
    }
    std::vector<thx_thu_record> rec(n);
    s = thx_thu_read(thuPath, rec.data(), n, &n);
    if (s) return s;
    const std::string prefix = parPrefix ? parPrefix : "";
    std::vector<float> images, attr(7 * (size_t)count);
    int N = 0;
    // stacks opened once each; images read in order (Optimiser::initImg,
    // src/Optimiser.cpp:4620-4660: "index@stack" with a 1-based slice index,
    // or a single-image file, both under parPrefix)
    std::map<std::string, std::pair<std::unique_ptr<File>, MrcHeader>> open;
    for (int l = 0; l < count; l++) {
        const thx_thu_record& r = rec[first + l];
        std::string name = r.particle, file;
        int slice = 0;
        const size_t at = name.find('@');
        if (at == std::string::npos) {
            file = prefix + name;
        } else {
            slice = std::atoi(name.substr(0, at).c_str()) - 1;
            file = prefix + name.substr(at + 1);
        }
        auto it = open.find(file);
        if (it == open.end()) {
            auto fp = std::make_unique<File>(file.c_str(), "rb");
            if (!fp->f) {
                set_error("load_thu: cannot open " + file);
                return THX_ERR_ARG;
            }
            MrcHeader h;
            if ((s = mrc_header(fp->f, h, file.c_str()))) return s;
            it = open.emplace(file, std::make_pair(std::move(fp), h)).first;
        }
        const MrcHeader& h = it->second.second;
        if (l == 0) {
            N = h.nx;
            images.resize((size_t)count * N * N);
        }
        if (h.nx != N || h.ny != N) {
            set_error("load_thu: image " + name + " is not " + std::to_string(N) + " x " +
                      std::to_string(N));
            return THX_ERR_ARG;
        }
        if ((s = mrc_slices(it->second.first->f, h, slice, 1, images.data() + (size_t)l * N * N,
                            file.c_str())))
            return s;
        std::memcpy(attr.data() + 7 * (size_t)l, r.ctf, 7 * sizeof(float));
    }
    s = thx_ctx_set_particle_images(ctx, images.data(), count, N, bgRadius, softMask, nullptr,
                                    nullptr, msOut);
    if (s) return s;
    return thx_ctx_set_ctf(ctx, attr.data(), count, pixelSize);
}

}  // extern "C"
