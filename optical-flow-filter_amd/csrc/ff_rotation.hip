// ff_rotation.hip -- rotational optical flow and Lax-Wendroff image
// propagation (gfx950): the image prediction path of
// RotationalFlowImagePredictor (src/gpu/rotation.cu:20-134), which shares
// the propagation idea of the filter with a second-order scheme.
//
// Arithmetic follows the reference kernels operation by operation. Where a
// reference expression lets its compiler fuse a product into the add or
// subtract that consumes it, the fusion is written out (fmaf / fma): the
// first product of `a*b - c*d` or `a*b + c*d`, and `z - h*d0 + k*d2`
// evaluated as two fused steps. The CPU oracle uses the same policy, so the
// GPU results are bitwise the oracle's.
//
//  * rotationalOpticalFlow_k (rotation_k.cu:17-43): the flow of a camera
//    rotating with angular velocity w, at pixel (x, y):
//      p = ((x - cx) / ax, (y - cy) / ay, 1), c = w x p,
//      flow = (ax c.x + (cx - x) c.z, ay c.y + (cy - y) c.z).
//  * LaxWendroffX_k / LaxWendroffY_k (propagation_k.h:49-132): one pass of
//    the Lax-Wendroff scheme along x or y with Courant number R = dt * flow:
//      out = z - R/2 (p - m) + R^2/2 (m - 2z + p),
//    neighbours read clamped at the image edges. On scalar images the
//    reference evaluates the update in double (its 0.5 is a double literal;
//    the Y pass's second difference too, through its 2.0), then rounds to
//    float; on 4-channel images every operation is float.
//
// Both are memory bound: 8 B of flow written per pixel; a scalar pass reads
// 8 B of flow and 4 B of image and writes 4 B (16 B/px), a 4-channel pass
// 40 B/px. One thread per pixel, a wave per 64-pixel row segment.
#include <hip/hip_runtime.h>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int h, int w) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY, 1); }

__global__ void __launch_bounds__(BX* BY) k_rotational_flow(Camera cam, float wx, float wy, float wz, DevImg flow) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= flow.w || y >= flow.h) return;
    const float fx = (float)x, fy = (float)y;
    // pixelToCameraCoordinates (camera_k.h:22-27)
    const float px = (fx - cam.centerX) / cam.alphaX;
    const float py = (fy - cam.centerY) / cam.alphaY;
    // cross(w, p) with p.z = 1 (math_k.h:21-26)
    const float cx = fmaf(wy, 1.0f, -(wz * py));
    const float cy = fmaf(wz, px, -(wx * 1.0f));
    const float cz = fmaf(wx, py, -(wy * px));
    const float u = fmaf(cam.alphaX, cx, (cam.centerX - fx) * cz);
    const float v = fmaf(cam.alphaY, cy, (cam.centerY - fy) * cz);
    *(float2*)((char*)flow.data + (size_t)y * flow.pitch + (size_t)x * 8) = make_float2(u, v);
}

template <int AXIS>
__device__ __forceinline__ float lw_scalar(float m, float z, float p, float R) {
    const float d0 = p - m;
    double d2;
    if constexpr (AXIS == 0) d2 = (double)((m - 2.0f * z) + p);  // float second difference
    else d2 = ((double)m - 2.0 * (double)z) + (double)p;         // double second difference
    const double h = 0.5 * (double)R;
    const double t = fma(-h, (double)d0, (double)z);
    return (float)fma(h * (double)R, d2, t);
}

__device__ __forceinline__ float lw_float(float m, float z, float p, float h, float k) {
    const float d0 = p - m;
    const float d2 = (m - 2.0f * z) + p;
    return fmaf(k, d2, fmaf(-h, d0, z));
}

// One Lax-Wendroff pass along AXIS (0: x, 1: y) of an image of C float
// channels (1 or 4), pixel-interleaved.
template <int AXIS, int C>
__global__ void __launch_bounds__(BX* BY) k_lax_wendroff(DevImg flow, DevImg in, DevImg out, float dt) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    const int h = out.h, w = out.w;
    if (x >= w || y >= h) return;
    const float2 f = *(const float2*)((const char*)flow.data + (size_t)y * flow.pitch + (size_t)x * 8);
    const float R = dt * (AXIS == 0 ? f.x : f.y);
    const int xm = AXIS == 0 ? max(x - 1, 0) : x, xp = AXIS == 0 ? min(x + 1, w - 1) : x;
    const int ym = AXIS == 1 ? max(y - 1, 0) : y, yp = AXIS == 1 ? min(y + 1, h - 1) : y;
    auto at = [&](int yy, int xx) {
        return (const float*)((const char*)in.data + (size_t)yy * in.pitch + (size_t)xx * C * 4);
    };
    float* o = (float*)((char*)out.data + (size_t)y * out.pitch + (size_t)x * C * 4);
    if constexpr (C == 1) {
        o[0] = lw_scalar<AXIS>(*at(ym, xm), *at(y, x), *at(yp, xp), R);
    } else {
        const float4 m = *(const float4*)at(ym, xm), z = *(const float4*)at(y, x), p = *(const float4*)at(yp, xp);
        const float hh = (float)(0.5 * (double)R);
        const float kk = (float)((0.5 * (double)R) * (double)R);
        *(float4*)o = make_float4(lw_float(m.x, z.x, p.x, hh, kk), lw_float(m.y, z.y, p.y, hh, kk),
                                  lw_float(m.z, z.z, p.z, hh, kk), lw_float(m.w, z.w, p.w, hh, kk));
    }
}

}  // namespace

void launch_rotational_flow(const Camera& cam, float wx, float wy, float wz, const DevImg& flow, hipStream_t s) {
    if (flow.h <= 0 || flow.w <= 0) return;
    ProfileScope ps("rotational_flow", s);
    hipLaunchKernelGGL(k_rotational_flow, grid_for(flow.h, flow.w), dim3(BX, BY), 0, s, cam, wx, wy, wz, flow);
}

void launch_lax_wendroff_pass(int axis, int channels, const DevImg& flow, const DevImg& in, const DevImg& out, float dt,
                              hipStream_t s) {
    if (out.h <= 0 || out.w <= 0) return;
    ProfileScope ps(axis == 0 ? "lax_wendroff_x" : "lax_wendroff_y", s);
    const dim3 g = grid_for(out.h, out.w), b(BX, BY);
    if (channels == 1) {
        if (axis == 0) hipLaunchKernelGGL((k_lax_wendroff<0, 1>), g, b, 0, s, flow, in, out, dt);
        else hipLaunchKernelGGL((k_lax_wendroff<1, 1>), g, b, 0, s, flow, in, out, dt);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_lax_wendroff<0, 4>), g, b, 0, s, flow, in, out, dt);
        else hipLaunchKernelGGL((k_lax_wendroff<1, 4>), g, b, 0, s, flow, in, out, dt);
    }
}

}  // namespace ffamd
