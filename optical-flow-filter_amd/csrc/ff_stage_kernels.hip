// ff_stage_kernels.hip -- per-stage HIP kernels for gfx950.
//
// One thread per output pixel, 64x4 blocks (a 64-lane wave spans one row
// segment so every load and store of a wave is one contiguous run).
// Texture reads of the reference are plain loads at clamped coordinates.
//
// Arithmetic follows the reference kernels operation by operation; the
// only fused multiply-adds are the explicit fmaf() calls, which stand for
// the `acc += a*b` / `x -= a*b` statements of the reference kernels (this
// file is compiled with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int h, int w) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY, 1); }

// ------------------------------------------------------------ image model

// imagePrefilter_k (imagemodel_k.cu:31-70) and imageModel_k
// (imagemodel_k.cu:76-133) evaluated in one kernel: the prefiltered values
// the model reads at (x+c, y) and (x, y+r) are recomputed from the image.
template <bool U8>
__global__ void __launch_bounds__(BX* BY) k_image_model(DevImg img, DevImg a0, DevImg grad) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    const int h = img.h, w = img.w;
    if (x >= w || y >= h) return;
    float dx = 0.0f, dy = 0.0f, sm = 0.0f;
#pragma unroll
    for (int c = -2; c <= 2; ++c) {
        // smooth_y at column x+c (component .x of the prefiltered pair)
        const int xc = clampi(x + c, 0, w - 1);
        float sy = 0.0f;
#pragma unroll
        for (int r = -2; r <= 2; ++r) sy = fmaf(smooth_w(r + 2), read_img<U8>(img, clampi(y + r, 0, h - 1), xc), sy);
        dx = fmaf(diff_w(c + 2), sy, dx);
        sm = fmaf(smooth_w(c + 2), sy, sm);
    }
#pragma unroll
    for (int r = -2; r <= 2; ++r) {
        // smooth_x at row y+r (component .y of the prefiltered pair)
        const int yr = clampi(y + r, 0, h - 1);
        float sx = 0.0f;
#pragma unroll
        for (int c = -2; c <= 2; ++c) sx = fmaf(smooth_w(c + 2), read_img<U8>(img, yr, clampi(x + c, 0, w - 1)), sx);
        dy = fmaf(diff_w(r + 2), sx, dy);
    }
    px<float>(a0, y, x) = sm;
    px<float2>(grad, y, x) = make_float2(dx, dy);
}

// -------------------------------------------------------------- misc

__global__ void __launch_bounds__(BX* BY) k_scale_f2(DevImg in, float s, DevImg out) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    const float2 v = px<float2>(in, y, x);
    px<float2>(out, y, x) = make_float2(s * v.x, s * v.y);
}

__global__ void __launch_bounds__(256) k_copy2d(DevImg src, DevImg dst, size_t row_bytes) {
    // 8- or 4-byte granules when every row start and the row length allow
    // it (k_copy2d_u64, k_copy2d_u32), bytes otherwise (launch_copy2d).
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (i >= row_bytes) return;
    const char* s = (const char*)src.data + (size_t)y * src.pitch;
    char* d = (char*)dst.data + (size_t)y * dst.pitch;
    d[i] = s[i];
}

__global__ void __launch_bounds__(256) k_copy2d_u64(DevImg src, DevImg dst, size_t row_words) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (i >= row_words) return;
    const uint64_t* s = (const uint64_t*)((const char*)src.data + (size_t)y * src.pitch);
    uint64_t* d = (uint64_t*)((char*)dst.data + (size_t)y * dst.pitch);
    d[i] = s[i];
}

__global__ void __launch_bounds__(256) k_copy2d_u32(DevImg src, DevImg dst, size_t row_words) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (i >= row_words) return;
    const uint32_t* s = (const uint32_t*)((const char*)src.data + (size_t)y * src.pitch);
    uint32_t* d = (uint32_t*)((char*)dst.data + (size_t)y * dst.pitch);
    d[i] = s[i];
}

// Many 2-D copies in one launch. A region is viewed as rows * (row_bytes /
// unit) accesses of `unit` bytes (16, or 4 / 1 when rows are not aligned
// for more); block b of a region copies kCopyPerBlock consecutive accesses,
// each thread two of them with both loads issued before the stores.
constexpr int kCopyMax = 64;
constexpr int kCopyPerBlock = 256 * 2;
struct CopyJob {
    const char* src;
    char* dst;
    long long src_pitch, dst_pitch;
    int per_row, total, first_block, unit;  // accesses per row and in all
};
// The job list travels as a kernel argument sized for the launch (one job,
// up to 8, up to kCopyMax): a single frame's copy does not pay for setting
// up 3 KB of arguments.
template <int M>
struct CopyBatch {
    int n;
    CopyJob job[M];
};

template <class T>
__device__ __forceinline__ const T* copy_src(const CopyJob& c, int i) {
    const int row = i / c.per_row, col = i - row * c.per_row;
    return reinterpret_cast<const T*>(c.src + row * c.src_pitch) + col;
}
template <class T>
__device__ __forceinline__ T* copy_dst(const CopyJob& c, int i) {
    const int row = i / c.per_row, col = i - row * c.per_row;
    return reinterpret_cast<T*>(c.dst + row * c.dst_pitch) + col;
}

// The two values stay in named registers: held in a two-element array they
// were placed in LDS by the compiler, with a wait for each load before the
// next one was issued (17.7 us for a 2048x1088 u8 frame instead of ~3 us).
template <class T>
__device__ __forceinline__ void copy_units(const CopyJob& c, int base) {
    const int i0 = base, i1 = base + 256;
    const bool ok0 = i0 < c.total, ok1 = i1 < c.total;
    T v0{}, v1{};
    if (ok0) v0 = *copy_src<T>(c, i0);
    if (ok1) v1 = *copy_src<T>(c, i1);
    if (ok0) *copy_dst<T>(c, i0) = v0;
    if (ok1) *copy_dst<T>(c, i1) = v1;
}

template <int M>
__global__ void __launch_bounds__(256) k_copy_regions(const CopyBatch<M> batch) {
    // the job whose block range holds this block: binary search over the
    // ascending first blocks (a linear scan costs a dependent scalar load
    // per job for the last regions' blocks)
    int lo = 0, hi = batch.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)blockIdx.x >= batch.job[mid].first_block) lo = mid;
        else hi = mid - 1;
    }
    const CopyJob& c = batch.job[lo];
    const int base = ((int)blockIdx.x - c.first_block) * kCopyPerBlock + (int)threadIdx.x;
    if (c.unit == 16) copy_units<uint4>(c, base);
    else if (c.unit == 4) copy_units<uint32_t>(c, base);
    else copy_units<uint8_t>(c, base);
}

// ------------------------------------------------------------ propagation

// flowPropagate{X,Y}_k and flowPropagatePayload{X,Y}_k
// (propagation_k.cu:16-334).
template <int AXIS, bool PAYLOAD>
__global__ void __launch_bounds__(BX* BY) k_propagate(DevImg fin, DevImg fout, DevImg sin, DevImg sout,
                                                      DevImg vin, DevImg vout, float dt, int border) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    const int h = fin.h, w = fin.w;
    if (x >= w || y >= h) return;
    const int ym = AXIS == 1 ? clampi(y - 1, 0, h - 1) : y, yp = AXIS == 1 ? clampi(y + 1, 0, h - 1) : y;
    const int xm = AXIS == 0 ? clampi(x - 1, 0, w - 1) : x, xp = AXIS == 0 ? clampi(x + 1, 0, w - 1) : x;
    const float2 fm = px<float2>(fin, ym, xm), f0 = px<float2>(fin, y, x), fp = px<float2>(fin, yp, xp);
    const float um = AXIS == 0 ? fm.x : fm.y, up = AXIS == 0 ? fp.x : fp.y;
    const float Ud = (fabsf(up) - fabsf(um)) > 0.0f ? up : um;
    const float ndtUd = -(dt * Ud);
    const bool back = Ud >= 0.0f;
    const bool in = (x >= border && x < w - border) && (y >= border && y < h - border);
    float2 r;
    r.x = upwind(fm.x, f0.x, fp.x, back, ndtUd);
    r.y = upwind(fm.y, f0.y, fp.y, back, ndtUd);
    px<float2>(fout, y, x) = in ? r : f0;
    if constexpr (PAYLOAD) {
        const float sm = px<float>(sin, ym, xm), s0 = px<float>(sin, y, x), sp = px<float>(sin, yp, xp);
        const float rs = upwind(sm, s0, sp, back, ndtUd);
        px<float>(sout, y, x) = in ? rs : s0;
        const float2 vm = px<float2>(vin, ym, xm), v0 = px<float2>(vin, y, x), vp = px<float2>(vin, yp, xp);
        float2 rv;
        rv.x = upwind(vm.x, v0.x, vp.x, back, ndtUd);
        rv.y = upwind(vm.y, v0.y, vp.y, back, ndtUd);
        px<float2>(vout, y, x) = in ? rv : v0;
    }
}

// ----------------------------------------------------------------- update

// flowUpdate_k (update_k.cu:15-79).
__global__ void __launch_bounds__(BX* BY) k_update(DevImg a0, DevImg grad, DevImg a0old, DevImg flow_old,
                                                   DevImg image_out, DevImg flow_out, float gamma,
                                                   float maxflow) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= a0.w || y >= a0.h) return;
    const float2 g = px<float2>(grad, y, x);
    const float a = px<float>(a0, y, x);
    const float Yt = px<float>(a0old, y, x) - a;
    const float2 f = px<float2>(flow_old, y, x);
    float ox, oy;
    solve2x2(g.x, g.y, Yt, f.x, f.y, gamma, ox, oy);
    ox = fmaxf(-maxflow, fminf(ox, maxflow));
    oy = fmaxf(-maxflow, fminf(oy, maxflow));
    ox = not_finite(ox) ? 0.0f : ox;
    oy = not_finite(oy) ? 0.0f : oy;
    px<float2>(flow_out, y, x) = make_float2(ox, oy);
    px<float>(image_out, y, x) = a;
}

// deltaFlowUpdate_k (update_k.cu:82-173).
__global__ void __launch_bounds__(BX* BY) k_delta_update(DevImg a0, DevImg grad, DevImg a0old, DevImg dflow_old,
                                                         DevImg upper, const int* __restrict__ xmap,
                                                         const int* __restrict__ ymap, DevImg image_out,
                                                         DevImg dflow_out, DevImg flow_out, float gamma,
                                                         float maxflow) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= a0.w || y >= a0.h) return;
    const float2 g = px<float2>(grad, y, x);
    const float a = px<float>(a0, y, x);
    const float Yt = px<float>(a0old, y, x) - a;
    const float2 d = px<float2>(dflow_old, y, x);
    float ox, oy;
    solve2x2(g.x, g.y, Yt, d.x, d.y, gamma, ox, oy);
    ox = not_finite(ox) ? 0.0f : ox;
    oy = not_finite(oy) ? 0.0f : oy;
    const float half = 0.5f * maxflow;
    ox = fmaxf(-half, fminf(ox, half));
    oy = fmaxf(-half, fminf(oy, half));
    const float2 fu = px<float2>(upper, ymap[y], xmap[x]);
    float fx = ox + 2.0f * fu.x, fy = oy + 2.0f * fu.y;
    fx = fmaxf(-maxflow, fminf(fx, maxflow));
    fy = fmaxf(-maxflow, fminf(fy, maxflow));
    px<float2>(dflow_out, y, x) = make_float2(ox, oy);
    px<float2>(flow_out, y, x) = make_float2(fx, fy);
    px<float>(image_out, y, x) = a;
}

// -------------------------------------------------------------- smoothing

// flowSmoothX_k / flowSmoothY_k (flowsmoothing_k.cu:21-87).
template <int AXIS>
__global__ void __launch_bounds__(BX* BY) k_smooth(DevImg in, DevImg out) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    const int h = in.h, w = in.w;
    if (x >= w || y >= h) return;
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int c = -2; c <= 2; ++c) {
        const int yy = AXIS == 1 ? clampi(y + c, 0, h - 1) : y;
        const int xx = AXIS == 0 ? clampi(x + c, 0, w - 1) : x;
        const float2 f = px<float2>(in, yy, xx);
        sx = fmaf(kBox, f.x, sx);
        sy = fmaf(kBox, f.y, sy);
    }
    px<float2>(out, y, x) = make_float2(sx, sy);
}

// ---------------------------------------------------------------- pyramid

// imageDown{X,Y}_uint8_k / imageDown{X,Y}_float_k (pyramid_k.cu:14-145).
template <int AXIS, bool U8>
__global__ void __launch_bounds__(BX* BY) k_down(DevImg in, DevImg out) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const int h = in.h, w = in.w;
    int ym = y, y0 = y, yp = y, xm = x, x0 = x, xp = x;
    if (AXIS == 0) { xm = clampi(2 * x - 1, 0, w - 1); x0 = clampi(2 * x, 0, w - 1); xp = clampi(2 * x + 1, 0, w - 1); }
    else { ym = clampi(2 * y - 1, 0, h - 1); y0 = clampi(2 * y, 0, h - 1); yp = clampi(2 * y + 1, 0, h - 1); }
    if constexpr (U8) {
        px<uint8_t>(out, y, x) = down_u8(px<uint8_t>(in, ym, xm), px<uint8_t>(in, y0, x0), px<uint8_t>(in, yp, xp));
    } else {
        const float tm = px<float>(in, ym, xm), t0 = px<float>(in, y0, x0), tp = px<float>(in, yp, xp);
        const float s = tm + tp;
        px<float>(out, y, x) = (float)(0.5 * (double)t0 + 0.25 * (double)s);
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_image_model(const DevImg& img, bool u8, const DevImg& a0, const DevImg& grad, hipStream_t s) {
    ProfileScope ps("image_model", s);
    if (img.h <= 0 || img.w <= 0) return;
    if (u8) hipLaunchKernelGGL(k_image_model<true>, grid_for(img.h, img.w), dim3(BX, BY), 0, s, img, a0, grad);
    else hipLaunchKernelGGL(k_image_model<false>, grid_for(img.h, img.w), dim3(BX, BY), 0, s, img, a0, grad);
}

void launch_scale_f2(const DevImg& in, float scalar, const DevImg& out, hipStream_t s) {
    ProfileScope ps("scale_f2", s);
    if (in.h <= 0 || in.w <= 0) return;
    hipLaunchKernelGGL(k_scale_f2, grid_for(in.h, in.w), dim3(BX, BY), 0, s, in, scalar, out);
}

void launch_propagate_pass(int axis, const DevImg& fin, const DevImg& fout, const DevImg* sin,
                           const DevImg* sout, const DevImg* vin, const DevImg* vout, float dt, int border,
                           hipStream_t s) {
    ProfileScope ps(sin ? (axis == 0 ? "propagate_payload_x" : "propagate_payload_y") : (axis == 0 ? "propagate_x" : "propagate_y"), s);
    if (fin.h <= 0 || fin.w <= 0) return;
    const dim3 g = grid_for(fin.h, fin.w), b(BX, BY);
    const DevImg none{};
    const bool pay = sin != nullptr;
    const DevImg si = pay ? *sin : none, so = pay ? *sout : none, vi = pay ? *vin : none, vo = pay ? *vout : none;
    if (axis == 0) {
        if (pay) hipLaunchKernelGGL((k_propagate<0, true>), g, b, 0, s, fin, fout, si, so, vi, vo, dt, border);
        else hipLaunchKernelGGL((k_propagate<0, false>), g, b, 0, s, fin, fout, si, so, vi, vo, dt, border);
    } else {
        if (pay) hipLaunchKernelGGL((k_propagate<1, true>), g, b, 0, s, fin, fout, si, so, vi, vo, dt, border);
        else hipLaunchKernelGGL((k_propagate<1, false>), g, b, 0, s, fin, fout, si, so, vi, vo, dt, border);
    }
}

void launch_update(const DevImg& a0, const DevImg& grad, const DevImg& a0old, const DevImg& flow_old,
                   const DevImg& image_out, const DevImg& flow_out, float gamma, float maxflow, hipStream_t s) {
    ProfileScope ps("update", s);
    if (a0.h <= 0 || a0.w <= 0) return;
    hipLaunchKernelGGL(k_update, grid_for(a0.h, a0.w), dim3(BX, BY), 0, s, a0, grad, a0old, flow_old, image_out,
                       flow_out, gamma, maxflow);
}

void launch_delta_update(const DevImg& a0, const DevImg& grad, const DevImg& a0old, const DevImg& dflow_old,
                         const DevImg& upper, const int* xmap, const int* ymap, const DevImg& image_out,
                         const DevImg& dflow_out, const DevImg& flow_out, float gamma, float maxflow,
                         hipStream_t s) {
    ProfileScope ps("delta_update", s);
    if (a0.h <= 0 || a0.w <= 0) return;
    hipLaunchKernelGGL(k_delta_update, grid_for(a0.h, a0.w), dim3(BX, BY), 0, s, a0, grad, a0old, dflow_old, upper,
                       xmap, ymap, image_out, dflow_out, flow_out, gamma, maxflow);
}

void launch_smooth_pass(int axis, const DevImg& in, const DevImg& out, hipStream_t s) {
    ProfileScope ps(axis == 0 ? "smooth_x" : "smooth_y", s);
    if (in.h <= 0 || in.w <= 0) return;
    if (axis == 0) hipLaunchKernelGGL(k_smooth<0>, grid_for(in.h, in.w), dim3(BX, BY), 0, s, in, out);
    else hipLaunchKernelGGL(k_smooth<1>, grid_for(in.h, in.w), dim3(BX, BY), 0, s, in, out);
}

void launch_down_pass(int axis, bool u8, const DevImg& in, const DevImg& out, hipStream_t s) {
    ProfileScope ps(axis == 0 ? "down_x" : "down_y", s);
    if (out.h <= 0 || out.w <= 0) return;
    const dim3 g = grid_for(out.h, out.w), b(BX, BY);
    if (axis == 0) {
        if (u8) hipLaunchKernelGGL((k_down<0, true>), g, b, 0, s, in, out);
        else hipLaunchKernelGGL((k_down<0, false>), g, b, 0, s, in, out);
    } else {
        if (u8) hipLaunchKernelGGL((k_down<1, true>), g, b, 0, s, in, out);
        else hipLaunchKernelGGL((k_down<1, false>), g, b, 0, s, in, out);
    }
}

namespace {
template <int M>
void launch_copy_batch(const CopyRegion* regions, int n, hipStream_t s) {
    CopyBatch<M> b{};
    int blocks = 0;
    for (int i = 0; i < n; ++i) {
        const CopyRegion& r = regions[i];
        if (r.rows <= 0 || r.row_bytes == 0) continue;
        const uintptr_t a = (uintptr_t)r.src | (uintptr_t)r.dst | r.src_pitch | r.dst_pitch | r.row_bytes;
        CopyJob& c = b.job[b.n++];
        c.src = (const char*)r.src;
        c.dst = (char*)r.dst;
        c.src_pitch = (long long)r.src_pitch;
        c.dst_pitch = (long long)r.dst_pitch;
        c.unit = (a % 16 == 0) ? 16 : ((a % 4 == 0) ? 4 : 1);
        c.per_row = (int)(r.row_bytes / c.unit);
        c.total = c.per_row * r.rows;
        c.first_block = blocks;
        blocks += (c.total + kCopyPerBlock - 1) / kCopyPerBlock;
    }
    if (!blocks) return;
    ProfileScope ps("copy_regions", s);
    hipLaunchKernelGGL(k_copy_regions<M>, dim3(blocks), dim3(256), 0, s, b);
}
}  // namespace

// Largest band of rows one copy job takes, in bytes. A job counts its
// accesses in 32-bit integers, and a launch of kCopyMax jobs of one-byte
// accesses (2^17 blocks of 256 threads each) stays below the 2^32
// work-items of a grid dimension; larger regions are split into bands.
constexpr size_t kCopyBandBytes = size_t(1) << 26;

void launch_copy_regions(const CopyRegion* regions, int n, hipStream_t s) {
    std::vector<CopyRegion> bands;
    bool split = false;
    for (int i = 0; i < n; ++i)
        split = split || (regions[i].rows > 0 && (size_t)regions[i].rows * regions[i].row_bytes > kCopyBandBytes);
    if (split) {
        for (int i = 0; i < n; ++i) {
            const CopyRegion& r = regions[i];
            if (r.rows <= 0 || r.row_bytes == 0) continue;
            if (r.row_bytes > kCopyBandBytes)
                throw invalid_argument("image copy: rows of more than 64 MiB are not supported (" +
                                       std::to_string(r.row_bytes) + " bytes)");
            const int per = (int)(kCopyBandBytes / r.row_bytes);
            for (int r0 = 0; r0 < r.rows; r0 += per) {
                CopyRegion band = r;
                band.src = (const char*)r.src + (size_t)r0 * r.src_pitch;
                band.dst = (char*)r.dst + (size_t)r0 * r.dst_pitch;
                band.rows = per < r.rows - r0 ? per : r.rows - r0;
                bands.push_back(band);
            }
        }
        regions = bands.data();
        n = (int)bands.size();
    }
    for (int first = 0; first < n; first += kCopyMax) {
        const int m = n - first < kCopyMax ? n - first : kCopyMax;
        if (m == 1) launch_copy_batch<1>(regions + first, m, s);
        else if (m <= 8) launch_copy_batch<8>(regions + first, m, s);
        else launch_copy_batch<kCopyMax>(regions + first, m, s);
    }
}

void launch_copy2d(const DevImg& src, const DevImg& dst, size_t row_bytes, hipStream_t s) {
    ProfileScope ps("copy2d", s);
    if (src.h <= 0 || row_bytes == 0) return;
    auto aligned = [&](size_t a) {
        return row_bytes % a == 0 && src.pitch % a == 0 && dst.pitch % a == 0 && (uintptr_t)src.data % a == 0 &&
               (uintptr_t)dst.data % a == 0;
    };
    if (aligned(8)) {
        const size_t n = row_bytes / 8;
        hipLaunchKernelGGL(k_copy2d_u64, dim3((unsigned)((n + 255) / 256), src.h), dim3(256), 0, s, src, dst, n);
    } else if (aligned(4)) {
        const size_t n = row_bytes / 4;
        hipLaunchKernelGGL(k_copy2d_u32, dim3((unsigned)((n + 255) / 256), src.h), dim3(256), 0, s, src, dst, n);
    } else {
        hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((row_bytes + 255) / 256), src.h), dim3(256), 0, s, src, dst,
                           row_bytes);
    }
}

}  // namespace ffamd
