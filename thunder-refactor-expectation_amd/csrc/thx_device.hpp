// Device helpers shared by the HIP translation units: rotation columns from
// a quaternion, the trilinear Fourier-space sample of a half-complex volume
// and the CTF closed form.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace thx {

// Geometry helpers (device)
// rotate3D (src/Geometry/Euler.cpp:181-189): R = I + 2 q0 [q]x + 2 [q]x^2.
// Only the first two columns are needed because slices lie in z = 0.
__device__ inline void quat_cols01(const double* q, double m[3][2]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    // A = [[0,-z,y],[z,0,-x],[-y,x,0]]; (A^2)(a,b) = sum_c A(a,c) A(c,b)
    const double A[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 2; b++) {
            double aa = A[a][0] * A[0][b] + A[a][1] * A[1][b] + A[a][2] * A[2][b];
            m[a][b] = (a == b ? 1.0 : 0.0) + 2 * w * A[a][b] + 2 * aa;
        }
}

// Trilinear Fourier-space sample of the half-complex volume at (x, y, z) in
// padded-grid units (Volume::getByInterpolationFT, src/Image/Volume.cpp:
// 308-332; getFTHalf(w, x0), :481-553).  Mirror and conjugate when x < 0;
// corners wrap in y and z.  The caller guarantees |(x, y, z)| + 1 < V / 2.
__device__ inline float2 sample_trilinear(const float2* __restrict__ vol, int V,
                                          float x, float y, float z) {
    bool conj = false;
    if (x < 0) { x = -x; y = -y; z = -z; conj = true; }
    const int x0 = (int)floorf(x), y0 = (int)floorf(y), z0 = (int)floorf(z);
    const float xd = x - x0, yd = y - y0, zd = z - z0;
    const float v0[2] = {1 - xd, xd}, v1[2] = {1 - yd, yd}, v2[2] = {1 - zd, zd};
    const int nColFT = V / 2 + 1;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        int zz = z0 + k;
        zz = zz < 0 ? zz + V : zz;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int yy = y0 + j;
            yy = yy < 0 ? yy + V : yy;
            const size_t row = ((size_t)zz * V + yy) * nColFT + x0;
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const float w = v0[i] * v1[j] * v2[k];
                const float2 d = vol[row + i];
                re += d.x * w;
                im += d.y * w;
            }
        }
    }
    return conj ? make_float2(re, -im) : make_float2(re, im);
}

// CTF(RFLOAT* dst, ...) closed form, src/CTF.cpp:109-145, in float (RFLOAT):
// u = |(i, j)| / (N pixelSize), angle = atan2(j, i) - theta, defocus =
// -(dU + dV + (dU - dV) cos 2 angle) / 2, ki = K1 defocus u^2 + K2 u^4 - phase
// shift, ctf = -sqrt(1 - A^2) sin ki + A cos ki, K1 = pi lambda,
// K2 = pi / 2 Cs lambda^3.
__device__ inline float ctf_value(int ic, int ir, float pixelSize, float lambda,
                                  float dU, float dV, float theta, float Cs,
                                  float ampC, int N, float phaseShift = 0.f) {
    const float w1 = sqrtf(1 - ampC * ampC), w2 = ampC;
    const float K1 = (float)M_PI * lambda;
    const float K2 = (float)(M_PI / 2 * (double)Cs * ((double)lambda * lambda * lambda));
    const float u = (float)hypot((double)(ic / (pixelSize * N)), (double)(ir / (pixelSize * N)));
    const float angle = (float)(atan2((double)ir, (double)ic) - theta);
    const float defocus = -(dU + dV + (dU - dV) * cosf(2 * angle)) / 2;
    const float u2 = u * u;
    const float ki = K1 * defocus * u2 + K2 * (u2 * u2) - phaseShift;
    return -w1 * sinf(ki) + w2 * cosf(ki);
}

}  // namespace thx
