// ff_display.hip -- colour coding of a flow field for display (gfx950).
//
// flowToColor_k (src/gpu/device/display_k.cu:14-43): each flow vector is
// normalised into the unit square, ((u + m) / (2m), (v + m) / (2m)) with m
// the stage's maxflow, and the colour wheel (a 256 x 256 RGBA table) is
// read at that point with the addressing of the reference's texture
// (normalised coordinates, point sampling, clamp): texel
// min(floor(c * n), n - 1) along each axis, 0 for c < 0. A NaN coordinate
// (NaN flow), which the texture unit's behaviour leaves open, reads texel 0.
//
// Memory-bound gather: 8 bytes of flow in and 4 bytes of colour out per
// pixel, the wheel (256 KiB) served from L2. Each thread handles four
// consecutive pixels of a row (two 16-byte loads, one 16-byte store).
#include <hip/hip_runtime.h>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int kLanes = 64, kRows = 4, kPx = 4;

// Texel index of normalised coordinate c on an axis of n texels.
__device__ __forceinline__ int wheel_index(float c, int n) {
    // fmaxf turns NaN and negative values into 0; the upper clamp comes
    // before the conversion so that it never overflows
    const float s = fminf(fmaxf(c * (float)n, 0.0f), (float)(n - 1));
    return (int)floorf(s);
}

__device__ __forceinline__ unsigned wheel_texel(const DevImg& wheel, float u, float v, float m, float den) {
    const int ix = wheel_index((u + m) / den, wheel.w);
    const int iy = wheel_index((v + m) / den, wheel.h);
    return *(const unsigned*)((const char*)wheel.data + (size_t)iy * wheel.pitch + (size_t)ix * 4);
}

template <bool VEC>
__global__ void __launch_bounds__(kLanes* kRows) k_flow_to_color(DevImg flow, DevImg wheel, float maxflow,
                                                                 DevImg color) {
    const int y = blockIdx.y * kRows + threadIdx.y;
    const int x0 = (blockIdx.x * kLanes + threadIdx.x) * kPx;
    if (y >= flow.h || x0 >= flow.w) return;
    const float2* f = (const float2*)((const char*)flow.data + (size_t)y * flow.pitch);
    unsigned* c = (unsigned*)((char*)color.data + (size_t)y * color.pitch);
    const float den = 2.0f * maxflow;
    if (VEC && x0 + kPx <= flow.w) {
        const float4 a = *(const float4*)(f + x0);
        const float4 b = *(const float4*)(f + x0 + 2);
        uint4 o;
        o.x = wheel_texel(wheel, a.x, a.y, maxflow, den);
        o.y = wheel_texel(wheel, a.z, a.w, maxflow, den);
        o.z = wheel_texel(wheel, b.x, b.y, maxflow, den);
        o.w = wheel_texel(wheel, b.z, b.w, maxflow, den);
        *(uint4*)(c + x0) = o;
        return;
    }
    for (int k = 0; k < kPx && x0 + k < flow.w; ++k) {
        const float2 v = f[x0 + k];
        c[x0 + k] = wheel_texel(wheel, v.x, v.y, maxflow, den);
    }
}

}  // namespace

void launch_flow_to_color(const DevImg& flow, const DevImg& wheel, float maxflow, const DevImg& color, hipStream_t s) {
    if (flow.h <= 0 || flow.w <= 0) return;
    ProfileScope ps("flow_to_color", s);
    const dim3 grid((flow.w + kLanes * kPx - 1) / (kLanes * kPx), (flow.h + kRows - 1) / kRows);
    // 16-byte accesses need 16-byte aligned rows
    const bool vec = ((uintptr_t)flow.data | flow.pitch | (uintptr_t)color.data | color.pitch) % 16 == 0;
    if (vec)
        hipLaunchKernelGGL(k_flow_to_color<true>, grid, dim3(kLanes, kRows), 0, s, flow, wheel, maxflow, color);
    else
        hipLaunchKernelGGL(k_flow_to_color<false>, grid, dim3(kLanes, kRows), 0, s, flow, wheel, maxflow, color);
}

}  // namespace ffamd
