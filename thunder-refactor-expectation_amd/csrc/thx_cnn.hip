// Learned pose proposal: a small CNN that maps a particle image to an
// angular-central-Gaussian (ACG) mixture over rotations and a Gaussian over
// shifts, with hand-written forward, backward and Adam kernels (FP32).
//
// Input: the particle's Fourier coefficients with 0 <= iCol < 32 and
// -32 <= iRow < 32 (the half plane of the pixel set; each stands for itself
// and its conjugate), phase-flipped by sign(CTF) and inverse-transformed to a
// 64 x 64 real image, normalised to zero mean and unit variance.
// Network (NCHW):
//   conv 1->16 5x5/2 (64->32) ReLU, conv 16->32 3x3/2 (->16) ReLU,
//   conv 32->64 3x3/2 (->8) ReLU, conv 64->128 3x3/2 (->4) ReLU,
//   fc 2048->384 ReLU, fc 384->NOUT (NOUT = 6 C + 3).
// Output per component c: raw quaternion (4) -> m_c = raw / |raw|,
//   a -> k_c = exp(-(1 + 6 sigmoid(a))), logit -> pi_c = softmax;
// then shift mean (2) and b -> sigma = exp(b) clipped to [0.2, 20].
// Loss: -log sum_c pi_c ACG_c(q) - log N(t; mu, sigma^2 I), with
//   ACG_c(q) = k^{-3/2} ((m.q)^2 + (1 - (m.q)^2) / k)^{-2} / (2 pi^2).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "thx_internal.hpp"

namespace thx {

constexpr int kImg = 64;
constexpr int kMaxBatch = 512;

struct ConvShape {
    int ci, hi, co, ho, ks, st, pd;
};
static const ConvShape kConv[4] = {
    {1, 64, 16, 32, 5, 2, 2}, {16, 32, 32, 16, 3, 2, 1}, {32, 16, 64, 8, 3, 2, 1},
    {64, 8, 128, 4, 3, 2, 1}};
constexpr int kFcIn = 128 * 4 * 4, kFcHid = 384;

// ---------------------------------------------------------------------------
// Images: 64 x 64 inverse DFT of the phase-flipped coefficients with
// |iCol|, |iRow| < 32 (Hermitian pairs included), then normalisation.
// One block per particle; the twiddle exp(2 pi i n / 64) is a table.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_images(const float2* __restrict__ dat,
                                                const float* __restrict__ ctf,
                                                const int* __restrict__ iCol,
                                                const int* __restrict__ iRow, int K, int P,
                                                float* __restrict__ img) {
    const int p = blockIdx.x;
    __shared__ float2 coef[2048];
    __shared__ int ci[2048], cj[2048];
    __shared__ float tw[2][64];
    __shared__ int nc;
    __shared__ float red[2][256];
    if (threadIdx.x < 64) {
        float s, c;
        sincospif(threadIdx.x / 32.f, &s, &c);
        tw[0][threadIdx.x] = c;
        tw[1][threadIdx.x] = s;
    }
    if (threadIdx.x == 0) nc = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const int i = iCol[k], j = iRow[k];
        if (i < 32 && j >= -32 && j < 32 && !(i == 0 && j == 0)) {
            const int slot = atomicAdd(&nc, 1);
            if (slot < 2048) {
                const float2 x = dat[(size_t)p * K + k];
                const float sg = ctf[(size_t)p * K + k] >= 0.f ? 1.f : -1.f;
                // a half-plane coefficient stands for itself and its conjugate
                const float f = 2.f;
                coef[slot] = make_float2(f * sg * x.x, f * sg * x.y);
                ci[slot] = i;
                cj[slot] = j;
            }
        }
    }
    __syncthreads();
    const int n = min(nc, 2048);
    float v[16];
    float s1 = 0.f, s2 = 0.f;
    for (int q = 0; q < 16; q++) {
        const int pix = threadIdx.x + 256 * q;
        const int y = pix >> 6, x = pix & 63;
        float acc = 0.f;
        for (int c = 0; c < n; c++) {
            const int ph = (ci[c] * x + cj[c] * y) & 63;
            acc += coef[c].x * tw[0][ph] - coef[c].y * tw[1][ph];
        }
        v[q] = acc;
        s1 += acc;
        s2 += acc * acc;
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    const float mean = red[0][0] / 4096.f;
    const float var = fmaxf(red[1][0] / 4096.f - mean * mean, 1e-20f);
    const float inv = rsqrtf(var);
    for (int q = 0; q < 16; q++) {
        const int pix = threadIdx.x + 256 * q;
        img[(size_t)p * 4096 + pix] = (v[q] - mean) * inv;
    }
}

// ---------------------------------------------------------------------------
// Layers
// ---------------------------------------------------------------------------
__global__ void k_conv_fwd(const float* __restrict__ in, const float* __restrict__ w,
                           const float* __restrict__ b, float* __restrict__ out, int B, int Ci,
                           int Hi, int Co, int Ho, int KS, int S, int Pd) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)B * Co * Ho * Ho;
    if (idx >= n) return;
    const int ox = idx % Ho, oy = (idx / Ho) % Ho, co = (idx / ((size_t)Ho * Ho)) % Co;
    const int bb = idx / ((size_t)Ho * Ho * Co);
    float acc = b[co];
    for (int c = 0; c < Ci; c++) {
        const float* ip = in + ((size_t)bb * Ci + c) * Hi * Hi;
        const float* wp = w + ((size_t)co * Ci + c) * KS * KS;
        for (int ky = 0; ky < KS; ky++) {
            const int iy = oy * S - Pd + ky;
            if (iy < 0 || iy >= Hi) continue;
            for (int kx = 0; kx < KS; kx++) {
                const int ix = ox * S - Pd + kx;
                if (ix < 0 || ix >= Hi) continue;
                acc += wp[ky * KS + kx] * ip[iy * Hi + ix];
            }
        }
    }
    out[idx] = fmaxf(acc, 0.f);   // every conv is followed by ReLU
}

// gradient w.r.t. the input of a conv (g = dL/d(pre-activation) of the output)
__global__ void k_conv_bwd_in(const float* __restrict__ g, const float* __restrict__ w,
                              float* __restrict__ din, int B, int Ci, int Hi, int Co, int Ho,
                              int KS, int S, int Pd) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)B * Ci * Hi * Hi;
    if (idx >= n) return;
    const int ix = idx % Hi, iy = (idx / Hi) % Hi, c = (idx / ((size_t)Hi * Hi)) % Ci;
    const int bb = idx / ((size_t)Hi * Hi * Ci);
    float acc = 0.f;
    for (int ky = 0; ky < KS; ky++) {
        const int ty = iy + Pd - ky;
        if (ty < 0 || ty % S) continue;
        const int oy = ty / S;
        if (oy >= Ho) continue;
        for (int kx = 0; kx < KS; kx++) {
            const int tx = ix + Pd - kx;
            if (tx < 0 || tx % S) continue;
            const int ox = tx / S;
            if (ox >= Ho) continue;
            for (int co = 0; co < Co; co++)
                acc += w[((size_t)co * Ci + c) * KS * KS + ky * KS + kx] *
                       g[(((size_t)bb * Co + co) * Ho + oy) * Ho + ox];
        }
    }
    din[idx] = acc;
}

// gradient w.r.t. weights: one block per weight, reduction over batch x output
__global__ __launch_bounds__(256) void k_conv_bwd_w(const float* __restrict__ g,
                                                    const float* __restrict__ in,
                                                    float* __restrict__ dw, int B, int Ci,
                                                    int Hi, int Co, int Ho, int KS, int S,
                                                    int Pd) {
    const int widx = blockIdx.x;
    const int kx = widx % KS, ky = (widx / KS) % KS, c = (widx / (KS * KS)) % Ci;
    const int co = widx / (KS * KS * Ci);
    const int per = Ho * Ho;
    float acc = 0.f;
    for (int t = threadIdx.x; t < B * per; t += blockDim.x) {
        const int bb = t / per, o = t % per, oy = o / Ho, ox = o % Ho;
        const int iy = oy * S - Pd + ky, ix = ox * S - Pd + kx;
        if (iy < 0 || iy >= Hi || ix < 0 || ix >= Hi) continue;
        acc += g[(((size_t)bb * Co + co) * Ho + oy) * Ho + ox] *
               in[(((size_t)bb * Ci + c) * Hi + iy) * Hi + ix];
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) dw[widx] = red[0];
}

// bias gradient: sum of g over batch and positions, one block per channel
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ g,
                                                   float* __restrict__ db, int B, int C, int HW) {
    const int c = blockIdx.x;
    float acc = 0.f;
    for (int t = threadIdx.x; t < B * HW; t += blockDim.x) {
        const int bb = t / HW, o = t % HW;
        acc += g[((size_t)bb * C + c) * HW + o];
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// g *= (act > 0): back through a ReLU whose output is act
__global__ void k_relu_back(float* __restrict__ g, const float* __restrict__ act, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && act[i] <= 0.f) g[i] = 0.f;
}

// y[b][o] = sum_i W[o][i] x[b][i] + bias[o], optional ReLU
__global__ void k_fc_fwd(const float* __restrict__ x, const float* __restrict__ W,
                         const float* __restrict__ bias, float* __restrict__ y, int B, int I,
                         int O, int relu) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int bb = blockIdx.y;
    if (o >= O) return;
    const float* xr = x + (size_t)bb * I;
    const float* wr = W + (size_t)o * I;
    float acc = bias[o];
    for (int i = 0; i < I; i++) acc += wr[i] * xr[i];
    y[(size_t)bb * O + o] = relu ? fmaxf(acc, 0.f) : acc;
}

// dx[b][i] = sum_o W[o][i] g[b][o]
__global__ void k_fc_bwd_x(const float* __restrict__ g, const float* __restrict__ W,
                           float* __restrict__ dx, int B, int I, int O) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int bb = blockIdx.y;
    if (i >= I) return;
    float acc = 0.f;
    for (int o = 0; o < O; o++) acc += W[(size_t)o * I + i] * g[(size_t)bb * O + o];
    dx[(size_t)bb * I + i] = acc;
}

// dW[o][i] = sum_b g[b][o] x[b][i]; db[o] = sum_b g[b][o]
__global__ void k_fc_bwd_w(const float* __restrict__ g, const float* __restrict__ x,
                           float* __restrict__ dW, float* __restrict__ db, int B, int I, int O) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)O * I) return;
    const int o = idx / I, i = idx % I;
    float acc = 0.f;
    for (int bb = 0; bb < B; bb++) acc += g[(size_t)bb * O + o] * x[(size_t)bb * I + i];
    dW[idx] = acc;
    if (i == 0) {
        float s = 0.f;
        for (int bb = 0; bb < B; bb++) s += g[(size_t)bb * O + o];
        db[o] = s;
    }
}

// ---------------------------------------------------------------------------
// Output head: mixture parameters, loss and its gradient w.r.t. the outputs.
// ---------------------------------------------------------------------------
__device__ inline float sigm(float a) { return 1.f / (1.f + expf(-a)); }

// mix (double): C x (m0..m3, k, pi) + (mu_x, mu_y, sigma)
__global__ void k_head(const float* __restrict__ out, int B, int C, double* __restrict__ mix) {
    const int bb = blockIdx.x * blockDim.x + threadIdx.x;
    if (bb >= B) return;
    const int NOUT = 6 * C + 3;
    const float* o = out + (size_t)bb * NOUT;
    double* m = mix + (size_t)bb * NOUT;
    float mx = -INFINITY;
    for (int c = 0; c < C; c++) mx = fmaxf(mx, o[6 * c + 5]);
    double zs = 0.0;
    for (int c = 0; c < C; c++) zs += exp((double)o[6 * c + 5] - mx);
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        for (int a = 0; a < 4; a++) m[6 * c + a] = r[a] / n;
        m[6 * c + 4] = exp(-(1.0 + 6.0 * sigm(r[4])));
        m[6 * c + 5] = exp((double)r[5] - mx) / zs;
    }
    m[6 * C] = o[6 * C];
    m[6 * C + 1] = o[6 * C + 1];
    m[6 * C + 2] = fmin(20.0, fmax(0.2, exp((double)o[6 * C + 2])));
}

// loss[b] and dL/dout[b] for targets q (4, double) and t (2, double); the
// batch loss is the mean, so the gradient carries 1 / B.
__global__ void k_loss(const float* __restrict__ out, const double* __restrict__ qt,
                       const double* __restrict__ tt, int B, int C, float* __restrict__ loss,
                       float* __restrict__ gout) {
    const int bb = blockIdx.x * blockDim.x + threadIdx.x;
    if (bb >= B) return;
    const int NOUT = 6 * C + 3;
    const float* o = out + (size_t)bb * NOUT;
    float* g = gout + (size_t)bb * NOUT;
    const double q[4] = {qt[4 * bb], qt[4 * bb + 1], qt[4 * bb + 2], qt[4 * bb + 3]};
    double lmax = -INFINITY;
    double lc[16], logpi[16], dlogA_dd[16], dlogA_dk[16], dd[16], kc[16];
    double mxl = -INFINITY;
    for (int c = 0; c < C; c++) mxl = fmax(mxl, (double)o[6 * c + 5]);
    double zs = 0.0;
    for (int c = 0; c < C; c++) zs += exp((double)o[6 * c + 5] - mxl);
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        double d = 0.0;
        for (int a = 0; a < 4; a++) d += r[a] / n * q[a];
        const double k = exp(-(1.0 + 6.0 * (double)sigm(r[4])));
        const double s = d * d + (1.0 - d * d) / k;
        const double logA = -log(2.0 * M_PI * M_PI) - 1.5 * log(k) - 2.0 * log(s);
        logpi[c] = (double)r[5] - mxl - log(zs);
        lc[c] = logpi[c] + logA;
        lmax = fmax(lmax, lc[c]);
        dd[c] = d;
        kc[c] = k;
        dlogA_dd[c] = -4.0 * d * (1.0 - 1.0 / k) / s;
        dlogA_dk[c] = -1.5 / k + 2.0 * (1.0 - d * d) / (k * k * s);
    }
    double se = 0.0;
    for (int c = 0; c < C; c++) se += exp(lc[c] - lmax);
    const double lrot = -(lmax + log(se));
    const double invB = 1.0 / B;
    for (int c = 0; c < C; c++) {
        const float* r = o + 6 * c;
        const double gam = exp(lc[c] - lmax) / se;   // responsibility
        const double n = sqrt((double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2] +
                              (double)r[3] * r[3]) + 1e-12;
        // dL/dlogA_c = -gam; d (m.q) / d raw = (q - d m) / n
        for (int a = 0; a < 4; a++) {
            const double ma = r[a] / n;
            g[6 * c + a] = (float)(invB * (-gam) * dlogA_dd[c] * (q[a] - dd[c] * ma) / n);
        }
        const double sg = sigm(r[4]);
        const double dk_da = -kc[c] * 6.0 * sg * (1.0 - sg);
        g[6 * c + 4] = (float)(invB * (-gam) * dlogA_dk[c] * dk_da);
        g[6 * c + 5] = (float)(invB * (exp(logpi[c]) - gam));
    }
    // shift: -log N(t; mu, sigma^2 I) = log(2 pi) + 2 b + |t - mu|^2 / (2 sigma^2)
    const double bs = o[6 * C + 2];
    const double sig = fmin(20.0, fmax(0.2, exp(bs)));
    const bool clipped = exp(bs) < 0.2 || exp(bs) > 20.0;
    const double dx = tt[2 * bb] - o[6 * C], dy = tt[2 * bb + 1] - o[6 * C + 1];
    const double r2 = dx * dx + dy * dy;
    const double lt = log(2.0 * M_PI) + 2.0 * log(sig) + r2 / (2.0 * sig * sig);
    g[6 * C] = (float)(invB * (-dx) / (sig * sig));
    g[6 * C + 1] = (float)(invB * (-dy) / (sig * sig));
    g[6 * C + 2] = clipped ? 0.f : (float)(invB * (2.0 - r2 / (sig * sig)));
    loss[bb] = (float)(lrot + lt);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_init_params(float* __restrict__ p, size_t n, float scale, uint64_t seed,
                              uint32_t stream) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[4];
    Philox::draw(seed, (uint32_t)i, (uint32_t)(i >> 32), stream, w);
    const double u1 = Philox::u01(w[0], w[1]), u2 = Philox::u01(w[2], w[3]);
    p[i] = (float)(scale * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
}

}  // namespace thx

// ---------------------------------------------------------------------------
// Network object and C ABI
// ---------------------------------------------------------------------------
using namespace thx;

struct thx_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    int C = 8, NOUT = 51;
    size_t off[13] = {};     // parameter offsets: w1 b1 w2 b2 w3 b3 w4 b4 W5 b5 W6 b6, total
    size_t nparam = 0;
    DevBuf params, grads, adam_m, adam_v;
    DevBuf act[7];           // input, conv1..4 outputs, fc1 output, head output
    DevBuf gact[7];
    DevBuf loss, qt, tt, mix, imgs;
    long step = 0;
};

static size_t conv_w(int l) {
    return (size_t)kConv[l].co * kConv[l].ci * kConv[l].ks * kConv[l].ks;
}

static int cnn_alloc(thx_cnn* n) {
    size_t o = 0;
    for (int l = 0; l < 4; l++) {
        n->off[2 * l] = o;
        o += conv_w(l);
        n->off[2 * l + 1] = o;
        o += kConv[l].co;
    }
    n->off[8] = o;
    o += (size_t)kFcHid * kFcIn;
    n->off[9] = o;
    o += kFcHid;
    n->off[10] = o;
    o += (size_t)n->NOUT * kFcHid;
    n->off[11] = o;
    o += n->NOUT;
    n->off[12] = o;
    n->nparam = o;
    THX_TRY(n->params.ensure(sizeof(float) * o));
    THX_TRY(n->grads.ensure(sizeof(float) * o));
    THX_TRY(n->adam_m.ensure(sizeof(float) * o));
    THX_TRY(n->adam_v.ensure(sizeof(float) * o));
    const size_t sz[7] = {(size_t)kImg * kImg, 16 * 32 * 32, 32 * 16 * 16, 64 * 8 * 8,
                          128 * 4 * 4, kFcHid, (size_t)n->NOUT};
    for (int i = 0; i < 7; i++) {
        THX_TRY(n->act[i].ensure(sizeof(float) * sz[i] * kMaxBatch));
        THX_TRY(n->gact[i].ensure(sizeof(float) * sz[i] * kMaxBatch));
    }
    THX_TRY(n->loss.ensure(sizeof(float) * kMaxBatch));
    THX_TRY(n->qt.ensure(sizeof(double) * 4 * kMaxBatch));
    THX_TRY(n->tt.ensure(sizeof(double) * 2 * kMaxBatch));
    THX_TRY(n->mix.ensure(sizeof(double) * n->NOUT * kMaxBatch));
    return THX_OK;
}

static float* P_(thx_cnn* n, int i) { return n->params.as<float>() + n->off[i]; }
static float* G_(thx_cnn* n, int i) { return n->grads.as<float>() + n->off[i]; }

static unsigned nblk(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// forward of B images already in act[0]
static int cnn_forward(thx_cnn* n, int B) {
    hipStream_t st = n->stream;
    for (int l = 0; l < 4; l++) {
        const ConvShape& c = kConv[l];
        const size_t no = (size_t)B * c.co * c.ho * c.ho;
        hipLaunchKernelGGL(k_conv_fwd, dim3(nblk(no)), dim3(256), 0, st, n->act[l].as<float>(),
                           P_(n, 2 * l), P_(n, 2 * l + 1), n->act[l + 1].as<float>(), B, c.ci, c.hi,
                           c.co, c.ho, c.ks, c.st, c.pd);
    }
    hipLaunchKernelGGL(k_fc_fwd, dim3(nblk(kFcHid, 128), B), dim3(128), 0, st,
                       n->act[4].as<float>(), P_(n, 8), P_(n, 9), n->act[5].as<float>(), B, kFcIn,
                       kFcHid, 1);
    hipLaunchKernelGGL(k_fc_fwd, dim3(nblk(n->NOUT, 64), B), dim3(64), 0, st,
                       n->act[5].as<float>(), P_(n, 10), P_(n, 11), n->act[6].as<float>(), B,
                       kFcHid, n->NOUT, 0);
    THX_HIP(hipGetLastError());
    return THX_OK;
}

// backward from gact[6] (dL/dout) to parameter gradients
static int cnn_backward(thx_cnn* n, int B) {
    hipStream_t st = n->stream;
    hipLaunchKernelGGL(k_fc_bwd_w, dim3(nblk((size_t)n->NOUT * kFcHid)), dim3(256), 0, st,
                       n->gact[6].as<float>(), n->act[5].as<float>(), G_(n, 10), G_(n, 11), B,
                       kFcHid, n->NOUT);
    hipLaunchKernelGGL(k_fc_bwd_x, dim3(nblk(kFcHid, 128), B), dim3(128), 0, st,
                       n->gact[6].as<float>(), P_(n, 10), n->gact[5].as<float>(), B, kFcHid,
                       n->NOUT);
    hipLaunchKernelGGL(k_relu_back, dim3(nblk((size_t)B * kFcHid)), dim3(256), 0, st,
                       n->gact[5].as<float>(), n->act[5].as<const float>(), (size_t)B * kFcHid);
    hipLaunchKernelGGL(k_fc_bwd_w, dim3(nblk((size_t)kFcHid * kFcIn)), dim3(256), 0, st,
                       n->gact[5].as<float>(), n->act[4].as<float>(), G_(n, 8), G_(n, 9), B,
                       kFcIn, kFcHid);
    hipLaunchKernelGGL(k_fc_bwd_x, dim3(nblk(kFcIn, 128), B), dim3(128), 0, st,
                       n->gact[5].as<float>(), P_(n, 8), n->gact[4].as<float>(), B, kFcIn,
                       kFcHid);
    for (int l = 3; l >= 0; l--) {
        const ConvShape& c = kConv[l];
        const size_t no = (size_t)B * c.co * c.ho * c.ho;
        hipLaunchKernelGGL(k_relu_back, dim3(nblk(no)), dim3(256), 0, st,
                           n->gact[l + 1].as<float>(), n->act[l + 1].as<const float>(), no);
        hipLaunchKernelGGL(k_conv_bwd_w, dim3((unsigned)conv_w(l)), dim3(256), 0, st,
                           n->gact[l + 1].as<float>(), n->act[l].as<float>(), G_(n, 2 * l), B,
                           c.ci, c.hi, c.co, c.ho, c.ks, c.st, c.pd);
        hipLaunchKernelGGL(k_bias_grad, dim3(c.co), dim3(256), 0, st, n->gact[l + 1].as<float>(),
                           G_(n, 2 * l + 1), B, c.co, c.ho * c.ho);
        if (l > 0) {
            const size_t ni = (size_t)B * c.ci * c.hi * c.hi;
            hipLaunchKernelGGL(k_conv_bwd_in, dim3(nblk(ni)), dim3(256), 0, st,
                               n->gact[l + 1].as<float>(), P_(n, 2 * l), n->gact[l].as<float>(),
                               B, c.ci, c.hi, c.co, c.ho, c.ks, c.st, c.pd);
        }
    }
    THX_HIP(hipGetLastError());
    return THX_OK;
}

extern "C" {

int thx_cnn_create(int device, int components, uint64_t seed, thx_cnn** out) {
    clear_error();
    THX_CHECK(out && components >= 1 && components <= 16, THX_ERR_ARG, "cnn_create: bad argument");
    thx_cnn* n = new thx_cnn();
    n->device = device;
    n->C = components;
    n->NOUT = 6 * components + 3;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        set_error("cnn_create: device or stream");
        return THX_ERR_HIP;
    }
    int s = cnn_alloc(n);
    if (s != THX_OK) { delete n; return s; }
    // He-normal weights, zero biases
    for (int l = 0; l < 4; l++) {
        const size_t nw = conv_w(l);
        const float sc = sqrtf(2.f / (kConv[l].ci * kConv[l].ks * kConv[l].ks));
        hipLaunchKernelGGL(k_init_params, dim3(nblk(nw)), dim3(256), 0, n->stream, P_(n, 2 * l),
                           nw, sc, seed, (uint32_t)(100 + l));
        (void)hipMemsetAsync(P_(n, 2 * l + 1), 0, sizeof(float) * kConv[l].co, n->stream);
    }
    hipLaunchKernelGGL(k_init_params, dim3(nblk((size_t)kFcHid * kFcIn)), dim3(256), 0, n->stream,
                       P_(n, 8), (size_t)kFcHid * kFcIn, sqrtf(2.f / kFcIn), seed, 110u);
    (void)hipMemsetAsync(P_(n, 9), 0, sizeof(float) * kFcHid, n->stream);
    hipLaunchKernelGGL(k_init_params, dim3(nblk((size_t)n->NOUT * kFcHid)), dim3(256), 0, n->stream,
                       P_(n, 10), (size_t)n->NOUT * kFcHid, 0.5f * sqrtf(1.f / kFcHid), seed, 111u);
    (void)hipMemsetAsync(P_(n, 11), 0, sizeof(float) * n->NOUT, n->stream);
    (void)hipMemsetAsync(n->adam_m.p, 0, sizeof(float) * n->nparam, n->stream);
    (void)hipMemsetAsync(n->adam_v.p, 0, sizeof(float) * n->nparam, n->stream);
    if (hipStreamSynchronize(n->stream) != hipSuccess) {
        set_error("cnn_create: init kernels failed");
        return THX_ERR_HIP;
    }
    *out = n;
    return THX_OK;
}

int thx_cnn_destroy(thx_cnn* n) {
    if (!n) return THX_OK;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->stream);
    DevBuf* b[] = {&n->params, &n->grads, &n->adam_m, &n->adam_v, &n->loss, &n->qt, &n->tt,
                   &n->mix, &n->imgs};
    for (DevBuf* x : b) x->release();
    for (int i = 0; i < 7; i++) {
        n->act[i].release();
        n->gact[i].release();
    }
    (void)hipStreamDestroy(n->stream);
    delete n;
    return THX_OK;
}

int thx_cnn_components(thx_cnn* n, int* C) {
    THX_CHECK(n && C, THX_ERR_ARG, "cnn_components: null argument");
    *C = n->C;
    return THX_OK;
}

int thx_cnn_num_params(thx_cnn* n, size_t* count) {
    THX_CHECK(n && count, THX_ERR_ARG, "cnn_num_params: null argument");
    *count = n->nparam;
    return THX_OK;
}

int thx_cnn_get_params(thx_cnn* n, float* p) {
    clear_error();
    THX_CHECK(n && p, THX_ERR_ARG, "cnn_get_params: null argument");
    THX_HIP(hipSetDevice(n->device));
    THX_HIP(hipStreamSynchronize(n->stream));
    THX_HIP(hipMemcpy(p, n->params.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return THX_OK;
}

int thx_cnn_set_params(thx_cnn* n, const float* p) {
    clear_error();
    THX_CHECK(n && p, THX_ERR_ARG, "cnn_set_params: null argument");
    THX_HIP(hipSetDevice(n->device));
    THX_HIP(hipStreamSynchronize(n->stream));
    THX_HIP(hipMemcpy(n->params.p, p, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    return THX_OK;
}

// Forward + loss (+ backward and parameter gradient when grad != NULL, or an
// Adam step when lr > 0).  images [B][64][64] float host, targets q [B][4],
// t [B][2] (double).  Returns the mean loss.
static int cnn_run(thx_cnn* n, const float* images, const float* dimages, int B,
                   const double* q, const double* t, double* lossOut, float* grad, float lr,
                   float* outHost) {
    THX_CHECK(n && B >= 1 && B <= kMaxBatch, THX_ERR_ARG, "cnn: batch must be in [1, 512]");
    THX_HIP(hipSetDevice(n->device));
    hipStream_t st = n->stream;
    if (images)
        THX_HIP(hipMemcpyAsync(n->act[0].p, images, sizeof(float) * B * 4096, hipMemcpyHostToDevice,
                               st));
    else
        THX_HIP(hipMemcpyAsync(n->act[0].p, dimages, sizeof(float) * B * 4096,
                               hipMemcpyDeviceToDevice, st));
    THX_TRY(cnn_forward(n, B));
    if (outHost) {
        THX_HIP(hipStreamSynchronize(st));
        THX_HIP(hipMemcpy(outHost, n->act[6].p, sizeof(float) * B * n->NOUT, hipMemcpyDeviceToHost));
    }
    if (!q) return THX_OK;
    THX_HIP(hipMemcpyAsync(n->qt.p, q, sizeof(double) * 4 * B, hipMemcpyHostToDevice, st));
    THX_HIP(hipMemcpyAsync(n->tt.p, t, sizeof(double) * 2 * B, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_loss, dim3(nblk(B, 64)), dim3(64), 0, st, n->act[6].as<const float>(),
                       n->qt.as<const double>(), n->tt.as<const double>(), B, n->C,
                       n->loss.as<float>(), n->gact[6].as<float>());
    THX_HIP(hipGetLastError());
    if (grad || lr > 0.f) {
        THX_TRY(cnn_backward(n, B));
        if (lr > 0.f) {
            n->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
            hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparam)), dim3(256), 0, st,
                               n->params.as<float>(), n->grads.as<const float>(),
                               n->adam_m.as<float>(), n->adam_v.as<float>(), n->nparam, lr, b1, b2,
                               1e-8f, c1, c2, 1e-5f);
            THX_HIP(hipGetLastError());
        }
    }
    std::vector<float> l(B);
    THX_HIP(hipStreamSynchronize(st));
    THX_HIP(hipMemcpy(l.data(), n->loss.p, sizeof(float) * B, hipMemcpyDeviceToHost));
    if (grad) THX_HIP(hipMemcpy(grad, n->grads.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int i = 0; i < B; i++) s += l[i];
    if (lossOut) *lossOut = s / B;
    return THX_OK;
}

int thx_cnn_forward(thx_cnn* n, const float* images, int B, float* out) {
    clear_error();
    THX_CHECK(images && out, THX_ERR_ARG, "cnn_forward: null argument");
    return cnn_run(n, images, nullptr, B, nullptr, nullptr, nullptr, nullptr, 0.f, out);
}

int thx_cnn_loss_grad(thx_cnn* n, const float* images, int B, const double* q, const double* t,
                      double* loss, float* grad) {
    clear_error();
    THX_CHECK(images && q && t && loss, THX_ERR_ARG, "cnn_loss_grad: null argument");
    return cnn_run(n, images, nullptr, B, q, t, loss, grad, 0.f, nullptr);
}

int thx_cnn_train_step(thx_cnn* n, const float* images, int B, const double* q, const double* t,
                       float lr, double* loss) {
    clear_error();
    THX_CHECK(images && q && t && lr > 0.f, THX_ERR_ARG, "cnn_train_step: bad argument");
    return cnn_run(n, images, nullptr, B, q, t, loss, nullptr, lr, nullptr);
}

// Images of the context's particles (device), optional host copy [P][64][64].
int thx_ctx_images(thx_ctx* c, float* host) {
    clear_error();
    THX_CHECK(c && c->P > 0 && c->K > 0, THX_ERR_STATE, "ctx_images: no particles");
    THX_HIP(hipSetDevice(c->device));
    THX_TRY(c->img.ensure(sizeof(float) * 4096 * (size_t)c->P));
    hipLaunchKernelGGL(k_images, dim3(c->P), dim3(256), 0, c->stream, c->dat.as<const float2>(),
                       c->ctf.as<const float>(), c->iCol.as<const int>(), c->iRow.as<const int>(),
                       c->K, c->P, c->img.as<float>());
    THX_HIP(hipGetLastError());
    if (host) {
        THX_HIP(hipStreamSynchronize(c->stream));
        THX_HIP(hipMemcpy(host, c->img.p, sizeof(float) * 4096 * (size_t)c->P, hipMemcpyDeviceToHost));
    }
    return THX_OK;
}

// Train on the context's particles (their images) with known poses: epochs
// passes of minibatches of size B in order, Adam with learning rate lr.
int thx_cnn_train_ctx(thx_cnn* n, thx_ctx* c, const double* q, const double* t, int B, int epochs,
                      float lr, double* lastLoss) {
    clear_error();
    THX_CHECK(n && c && q && t && B >= 1 && B <= kMaxBatch && epochs >= 1, THX_ERR_ARG,
              "cnn_train_ctx: bad argument");
    THX_TRY(thx_ctx_images(c, nullptr));
    THX_HIP(hipStreamSynchronize(c->stream));
    double l = 0.0;
    for (int e = 0; e < epochs; e++)
        for (int b0 = 0; b0 + B <= c->P; b0 += B)
            THX_TRY(cnn_run(n, nullptr, c->img.as<float>() + (size_t)b0 * 4096, B, q + 4 * (size_t)b0,
                            t + 2 * (size_t)b0, &l, nullptr, lr, nullptr));
    if (lastLoss) *lastLoss = l;
    return THX_OK;
}

// Proposal parameters for every particle of the context: mix[P][6C + 3].
int thx_cnn_propose_ctx(thx_cnn* n, thx_ctx* c, double* mixHost, double* mixDev) {
    clear_error();
    THX_CHECK(n && c && (mixHost || mixDev), THX_ERR_ARG, "cnn_propose_ctx: bad argument");
    THX_TRY(thx_ctx_images(c, nullptr));
    THX_HIP(hipStreamSynchronize(c->stream));
    for (int b0 = 0; b0 < c->P; b0 += kMaxBatch) {
        const int B = std::min(kMaxBatch, c->P - b0);
        THX_TRY(cnn_run(n, nullptr, c->img.as<float>() + (size_t)b0 * 4096, B, nullptr, nullptr,
                        nullptr, nullptr, 0.f, nullptr));
        hipLaunchKernelGGL(k_head, dim3(nblk(B, 64)), dim3(64), 0, n->stream,
                           n->act[6].as<const float>(), B, n->C, n->mix.as<double>());
        THX_HIP(hipGetLastError());
        THX_HIP(hipStreamSynchronize(n->stream));
        const size_t sz = sizeof(double) * B * n->NOUT;
        if (mixHost)
            THX_HIP(hipMemcpy(mixHost + (size_t)b0 * n->NOUT, n->mix.p, sz, hipMemcpyDeviceToHost));
        if (mixDev)
            THX_HIP(hipMemcpy(mixDev + (size_t)b0 * n->NOUT, n->mix.p, sz, hipMemcpyDeviceToDevice));
    }
    return THX_OK;
}

}  // extern "C"
