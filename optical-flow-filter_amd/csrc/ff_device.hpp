// ff_device.hpp -- per-pixel arithmetic shared by the stage kernels and the
// fused level kernels (device code only).
//
// The operation order is the reference kernels' (file:line cited per
// helper); the only fused multiply-adds are the explicit fmaf() calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ff_kernels.hpp"

namespace ffamd {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class T>
__device__ __forceinline__ T& px(const DevImg& im, int y, int x) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(im.data) + (size_t)y * im.pitch)[x];
}

// Binomial and derivative masks, imagemodel_k.cu:21-22 (exact in binary).
__device__ __forceinline__ float smooth_w(int i) {
    return (i == 0 || i == 4) ? 0.0625f : ((i == 1 || i == 3) ? 0.25f : 0.375f);
}
__device__ __forceinline__ float diff_w(int i) {
    return i == 0 ? -0.125f : (i == 1 ? -0.25f : (i == 2 ? 0.0f : (i == 3 ? 0.25f : 0.125f)));
}
// Box weight of the flow smoother, flowsmoothing_k.cu:19.
static constexpr float kBox = 0.2f;

// Normalized read of an 8-bit texel, defined as (float)v / 255.0f. Computed
// as a product with the rounded reciprocal plus one fused correction step;
// this equals the IEEE quotient for every v in 0..255 (checked exhaustively
// by the CPU test suite), at three instructions instead of a division.
__device__ __forceinline__ float unorm8(uint32_t v) {
    const float r = 1.0f / 255.0f;
    const float f = (float)v;
    const float q = f * r;
    const float e = fmaf(-q, 255.0f, f);
    return fmaf(e, r, q);
}

template <bool U8>
__device__ __forceinline__ float read_img(const DevImg& im, int y, int x) {
    if constexpr (U8) return unorm8(px<uint8_t>(im, y, x));
    else return px<float>(im, y, x);
}

// Upwind update of one component, propagation_k.cu:45-53:
//   z - (dt*Ud) * (Ud >= 0 ? z - m : p - z), with ndtUd = -(dt*Ud).
__device__ __forceinline__ float upwind(float m, float z, float p, bool back, float ndtUd) {
    const float d = back ? (z - m) : (p - z);
    return fmaf(ndtUd, d, z);
}

// Closed-form regularised 2x2 solve, update_k.cu:43-63 (and :114-134).
__device__ __forceinline__ void solve2x2(float gx, float gy, float Yt, float fx, float fy, float gamma, float& ox,
                                         float& oy) {
    const float ax2 = gx * gx;
    const float ay2 = gy * gy;
    const float N00 = gamma + ay2;
    const float N01 = -gx * gy;
    const float N11 = gamma + ax2;
    const float rdet = 1.0f / (gamma * ((gamma + ax2) + ay2));
    const float qx = fmaf(gamma, fx, gx * Yt);
    const float qy = fmaf(gamma, fy, gy * Yt);
    ox = fmaf(N00, qx, N01 * qy) * rdet;
    oy = fmaf(N01, qx, N11 * qy) * rdet;
}

__device__ __forceinline__ bool not_finite(float v) { return !__builtin_isfinite(v); }

// u8 pyramid tap, imageDown{X,Y}_uint8_k (pyramid_k.cu:38-43, 72-77):
// (uchar)(255 * (float)(0.5*(double)t0 + 0.25*(double)(tm + tp))) with
// t = v/255.0f. For all 2^24 input triples this equals the integer form
// below (checked exhaustively by the CPU test suite).
__device__ __forceinline__ uint8_t down_u8(uint32_t km, uint32_t k0, uint32_t kp) {
    return (uint8_t)((2u * k0 + km + kp) >> 2);
}

}  // namespace ffamd
