// ff_fused.hip -- one kernel per pyramid level and frame for gfx950.
//
// The kernel runs, for one output tile, the whole per-level recurrence of
// the reference (FlowFilter::compute, src/gpu/flowfilter.cu:125-153, and the
// DeltaFlowFilter update path, :429-503): N upwind propagation iterations of
// the previous smoothed flow (with the old brightness and delta flow as
// payloads on delta levels), the image model, the 2x2 update and K box
// smoothing iterations. State is read once and written once per frame;
// everything in between stays in registers.
//
// Geometry: a workgroup is NW waves stacked vertically; lane l of every wave
// owns buffer column l (64 columns), and each thread owns R consecutive rows
// of its wave's strip, so the buffer is 64 x (NW*R) pixels. The output tile
// is the buffer minus a halo h0 = max(N,2) + 2K on every side (the distance
// an error at the buffer edge travels inward: one pixel per propagation
// pass, two per smoothing pass, two for the image model's x taps).
//  - X neighbours come from DPP wave shifts (a lane at the buffer edge reads
//    itself; the resulting error stays inside the halo).
//  - Y neighbours are registers of the same thread; the rows above and below
//    a wave's strip are exchanged through LDS once per Y pass.
//  - Every buffer lies inside the image: buffers at the image edges are
//    clamped against the edge and their output range extends to it. A lane
//    without a DPP source keeps its own value and the first/last wave of a
//    workgroup uses its own first/last row for the missing neighbour rows,
//    which is exactly the clamp-to-edge read of the reference at an image
//    edge (and a bounded error inside the halo elsewhere). The fused path
//    therefore needs W >= 64 and H >= NW*R.
//
// Arithmetic per pixel is the reference's (same operations, order and
// rounding points as the stage kernels and the CPU oracle); results are
// bitwise identical to the stage-by-stage path.
#include <hip/hip_runtime.h>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int LANES = 64;

struct FusedParams {
    const void* img;
    const float2* flow_in;
    const float* a0_in;
    const float2* dflow_in;
    float2* flow_out;
    float* a0_out;
    float2* dflow_out;
    float2* flow_upd;  // may be null
    const float2* upper;
    const int* xmap;
    const int* ymap;
    // row pitches in bytes
    int img_pitch, flow_in_pitch, a0_in_pitch, dflow_in_pitch, flow_out_pitch, a0_out_pitch, dflow_out_pitch,
        flow_upd_pitch, upper_pitch;
    int W, H, N, K, border, h0, TW, TH, tiles_x, tiles_y;
    float dt, gamma, maxflow;
};

template <class T>
__device__ __forceinline__ T ld(const T* base, int pitch, int y, int x) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)y * pitch)[x];
}
template <class T>
__device__ __forceinline__ void st(T* base, int pitch, int y, int x, T v) {
    reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)y * pitch)[x] = v;
}

// DPP wave shifts: lane i receives lane i-1 (shr) or lane i+1 (shl); the lane
// without a source keeps its own value.
__device__ __forceinline__ float wshr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float wshl(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ float rdlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// 5-tap correlation along x in the order of the reference loops (c = -2..2):
// acc = 0; acc = fma(w[c], v(x+c), acc).
template <class Wf>
__device__ __forceinline__ float taps_x(float v, Wf w) {
    const float m1 = wshr(v), m2 = wshr(m1), p1 = wshl(v), p2 = wshl(p1);
    float a = 0.0f;
    a = fmaf(w(0), m2, a);
    a = fmaf(w(1), m1, a);
    a = fmaf(w(2), v, a);
    a = fmaf(w(3), p1, a);
    a = fmaf(w(4), p2, a);
    return a;
}

struct WSmooth {
    __device__ __forceinline__ float operator()(int i) const { return smooth_w(i); }
};
struct WDiff {
    __device__ __forceinline__ float operator()(int i) const { return diff_w(i); }
};
struct WBox {
    __device__ __forceinline__ float operator()(int) const { return kBox; }
};

// Map block ids so that the tiles of one XCD (blocks are dealt round-robin
// over the 8 XCDs) form a contiguous band of the image.
__device__ __forceinline__ int xcd_tile(int b, int nt) {
    const int q = nt / 8;
    if (b >= 8 * q) return b;
    return (b % 8) * q + b / 8;
}

// Per-thread view of the tile.
struct Geo {
    int lane, wave, x, xc, yw, h0;
    bool xin;      // column strictly inside the propagation border
    bool wave_in;  // every pixel of the wave strictly inside the border
};

// Registers of the level state: flow (fu, fv) and, on delta levels, the
// scalar (old brightness) and vector (delta flow) payloads.
template <bool DELTA, int R>
struct State {
    float fu[R], fv[R];
    float sa[DELTA ? R : 1], du[DELTA ? R : 1], dv[DELTA ? R : 1];
};

// One upwind X pass (flowPropagateX_k / flowPropagatePayloadX_k,
// propagation_k.cu:16-72, 134-232).
template <bool DELTA, bool BORDER, int R>
__device__ __forceinline__ void prop_x(State<DELTA, R>& s, const Geo& g, const FusedParams& p) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float zu = s.fu[r], zv = s.fv[r];
        const float mu = wshr(zu), pu = wshl(zu);
        const float Ud = (fabsf(pu) - fabsf(mu)) > 0.0f ? pu : mu;
        const float nd = -(p.dt * Ud);
        const bool back = Ud >= 0.0f;
        float ou = upwind(mu, zu, pu, back, nd);
        float ov = upwind(wshr(zv), zv, wshl(zv), back, nd);
        float os = 0.f, odu = 0.f, odv = 0.f;
        if constexpr (DELTA) {
            os = upwind(wshr(s.sa[r]), s.sa[r], wshl(s.sa[r]), back, nd);
            odu = upwind(wshr(s.du[r]), s.du[r], wshl(s.du[r]), back, nd);
            odv = upwind(wshr(s.dv[r]), s.dv[r], wshl(s.dv[r]), back, nd);
        }
        if constexpr (BORDER) {
            const int y = g.yw + r;
            const bool in = g.xin && y >= p.border && y < p.H - p.border;
            ou = in ? ou : zu;
            ov = in ? ov : zv;
            if constexpr (DELTA) {
                os = in ? os : s.sa[r];
                odu = in ? odu : s.du[r];
                odv = in ? odv : s.dv[r];
            }
        }
        s.fu[r] = ou;
        s.fv[r] = ov;
        if constexpr (DELTA) {
            s.sa[r] = os;
            s.du[r] = odu;
            s.dv[r] = odv;
        }
    }
}

// Rows above (a) and below (b) a wave's strip for one Y pass.
struct Edge5 {
    float u, v, s, du, dv;
};

// One upwind Y pass (flowPropagateY_k / flowPropagatePayloadY_k,
// propagation_k.cu:75-131, 235-334); rows top-down, m* keep the old values
// of row r-1.
template <bool DELTA, bool BORDER, int R>
__device__ __forceinline__ void prop_y(State<DELTA, R>& s, const Geo& g, const FusedParams& p, const Edge5& a,
                                       const Edge5& b) {
    float mu = a.u, mv = a.v, ms = a.s, mdu = a.du, mdv = a.dv;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float zu = s.fu[r], zv = s.fv[r];
        const float pu = r + 1 < R ? s.fu[r + 1] : b.u;
        const float pv = r + 1 < R ? s.fv[r + 1] : b.v;
        const float Vd = (fabsf(pv) - fabsf(mv)) > 0.0f ? pv : mv;
        const float nd = -(p.dt * Vd);
        const bool back = Vd >= 0.0f;
        float ou = upwind(mu, zu, pu, back, nd);
        float ov = upwind(mv, zv, pv, back, nd);
        mu = zu;
        mv = zv;
        bool in = true;
        if constexpr (BORDER) {
            const int y = g.yw + r;
            in = g.xin && y >= p.border && y < p.H - p.border;
            ou = in ? ou : zu;
            ov = in ? ov : zv;
        }
        s.fu[r] = ou;
        s.fv[r] = ov;
        if constexpr (DELTA) {
            const float zs = s.sa[r], zdu = s.du[r], zdv = s.dv[r];
            const float ps = r + 1 < R ? s.sa[r + 1] : b.s;
            const float pdu = r + 1 < R ? s.du[r + 1] : b.du;
            const float pdv = r + 1 < R ? s.dv[r + 1] : b.dv;
            float os = upwind(ms, zs, ps, back, nd);
            float odu = upwind(mdu, zdu, pdu, back, nd);
            float odv = upwind(mdv, zdv, pdv, back, nd);
            ms = zs;
            mdu = zdu;
            mdv = zdv;
            if constexpr (BORDER) {
                os = in ? os : zs;
                odu = in ? odu : zdu;
                odv = in ? odv : zdv;
            }
            s.sa[r] = os;
            s.du[r] = odu;
            s.dv[r] = odv;
        }
    }
}

template <bool DELTA, bool U8, int R, int NW>
__global__ void __launch_bounds__(NW* LANES) k_fused_level(FusedParams p) {
    // row exchange between the waves of the workgroup, double buffered
    __shared__ float xch[2][NW][10][LANES];

    Geo g;
    g.lane = threadIdx.x & (LANES - 1);
    g.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = xcd_tile(blockIdx.x, p.tiles_x * p.tiles_y);
    const int tyi = t / p.tiles_x, txi = t - tyi * p.tiles_x;
    // output range of the tile and the buffer around it, clamped into the image
    const int ox0 = txi * p.TW, ox1 = min(ox0 + p.TW, p.W);
    const int oy0 = tyi * p.TH, oy1 = min(oy0 + p.TH, p.H);
    const int x0 = clampi(ox0 - p.h0, 0, p.W - LANES);
    const int y0 = clampi(oy0 - p.h0, 0, p.H - NW * R);
    g.h0 = p.h0;
    g.x = x0 + g.lane;
    g.xc = g.x;
    g.yw = y0 + g.wave * R;  // image row of this thread's first row (wave-uniform)
    const int b = p.border;
    g.xin = g.x >= b && g.x < p.W - b;
    g.wave_in = (x0 >= b) && (x0 + LANES - 1 < p.W - b) && (g.yw >= b) && (g.yw + R - 1 < p.H - b);
    const int lane = g.lane, wave = g.wave;
    const bool out_x = g.x >= ox0 && g.x < ox1;
    auto out_row = [&](int r) { return g.yw + r >= oy0 && g.yw + r < oy1; };

    State<DELTA, R> s;

    // ------------------------------------------------------ load the state
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int yc = g.yw + r;
        const float2 f = ld(p.flow_in, p.flow_in_pitch, yc, g.xc);
        s.fu[r] = f.x;
        s.fv[r] = f.y;
        if constexpr (DELTA) {
            s.sa[r] = ld(p.a0_in, p.a0_in_pitch, yc, g.xc);
            const float2 d = ld(p.dflow_in, p.dflow_in_pitch, yc, g.xc);
            s.du[r] = d.x;
            s.dv[r] = d.y;
        }
    }

    int slot = 0;

    // ---------------------------------------------------------- propagation
    for (int it = 0; it < p.N; ++it) {
        if (g.wave_in) prop_x<DELTA, false, R>(s, g, p);
        else prop_x<DELTA, true, R>(s, g, p);
        {
            float* my = &xch[slot][wave][0][0];
            my[0 * LANES + lane] = s.fu[0];
            my[1 * LANES + lane] = s.fv[0];
            my[5 * LANES + lane] = s.fu[R - 1];
            my[6 * LANES + lane] = s.fv[R - 1];
            if constexpr (DELTA) {
                my[2 * LANES + lane] = s.sa[0];
                my[3 * LANES + lane] = s.du[0];
                my[4 * LANES + lane] = s.dv[0];
                my[7 * LANES + lane] = s.sa[R - 1];
                my[8 * LANES + lane] = s.du[R - 1];
                my[9 * LANES + lane] = s.dv[R - 1];
            }
        }
        __syncthreads();
        // missing neighbour rows (first/last wave): the own edge row, the
        // clamp-to-edge read at an image edge
        Edge5 ea{s.fu[0], s.fv[0], 0.f, 0.f, 0.f}, eb{s.fu[R - 1], s.fv[R - 1], 0.f, 0.f, 0.f};
        if constexpr (DELTA) {
            ea.s = s.sa[0];
            ea.du = s.du[0];
            ea.dv = s.dv[0];
            eb.s = s.sa[R - 1];
            eb.du = s.du[R - 1];
            eb.dv = s.dv[R - 1];
        }
        if (wave > 0) {
            const float* o = &xch[slot][wave - 1][0][0];
            ea.u = o[5 * LANES + lane];
            ea.v = o[6 * LANES + lane];
            if constexpr (DELTA) {
                ea.s = o[7 * LANES + lane];
                ea.du = o[8 * LANES + lane];
                ea.dv = o[9 * LANES + lane];
            }
        }
        if (wave < NW - 1) {
            const float* o = &xch[slot][wave + 1][0][0];
            eb.u = o[0 * LANES + lane];
            eb.v = o[1 * LANES + lane];
            if constexpr (DELTA) {
                eb.s = o[2 * LANES + lane];
                eb.du = o[3 * LANES + lane];
                eb.dv = o[4 * LANES + lane];
            }
        }
        slot ^= 1;
        if (g.wave_in) prop_y<DELTA, false, R>(s, g, p, ea, eb);
        else prop_y<DELTA, true, R>(s, g, p, ea, eb);
    }

    // ------------------------------------------------ image model + update
    // imagePrefilter_k / imageModel_k (imagemodel_k.cu:31-133), then
    // flowUpdate_k (update_k.cu:15-79) or deltaFlowUpdate_k (:82-173), one
    // row at a time with a 5-row sliding window of image rows (w*) and of
    // their X-smoothed values (x*), read with clamped coordinates.
    auto img_row = [&](int j) -> float {  // image row yw - 2 + j
        const int yc = clampi(g.yw - 2 + j, 0, p.H - 1);
        if constexpr (U8) return unorm8(ld(reinterpret_cast<const uint8_t*>(p.img), p.img_pitch, yc, g.xc));
        else return ld(reinterpret_cast<const float*>(p.img), p.img_pitch, yc, g.xc);
    };
    float w0 = img_row(0), w1 = img_row(1), w2 = img_row(2), w3 = img_row(3);
    float x0s = taps_x(w0, WSmooth{}), x1s = taps_x(w1, WSmooth{}), x2s = taps_x(w2, WSmooth{}),
          x3s = taps_x(w3, WSmooth{});
    const int xmx = DELTA ? p.xmap[g.xc] : 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float w4 = img_row(r + 4);
        const float x4s = taps_x(w4, WSmooth{});
        const int y = g.yw + r;
        const int yc = y;
        float sy = 0.0f;
        sy = fmaf(smooth_w(0), w0, sy);
        sy = fmaf(smooth_w(1), w1, sy);
        sy = fmaf(smooth_w(2), w2, sy);
        sy = fmaf(smooth_w(3), w3, sy);
        sy = fmaf(smooth_w(4), w4, sy);
        const float a0 = taps_x(sy, WSmooth{});
        const float gx = taps_x(sy, WDiff{});
        float gy = 0.0f;
        gy = fmaf(diff_w(0), x0s, gy);
        gy = fmaf(diff_w(1), x1s, gy);
        gy = fmaf(diff_w(2), x2s, gy);
        gy = fmaf(diff_w(3), x3s, gy);
        gy = fmaf(diff_w(4), x4s, gy);
        w0 = w1; w1 = w2; w2 = w3; w3 = w4;
        x0s = x1s; x1s = x2s; x2s = x3s; x3s = x4s;
        const bool out = out_x && out_row(r);
        float ox, oy;
        if constexpr (DELTA) {
            const float Yt = s.sa[r] - a0;
            solve2x2(gx, gy, Yt, s.du[r], s.dv[r], p.gamma, ox, oy);
            ox = not_finite(ox) ? 0.0f : ox;
            oy = not_finite(oy) ? 0.0f : oy;
            const float half = 0.5f * p.maxflow;
            ox = fmaxf(-half, fminf(ox, half));
            oy = fmaxf(-half, fminf(oy, half));
            const float2 up = ld(p.upper, p.upper_pitch, p.ymap[yc], xmx);
            float fx = ox + 2.0f * up.x, fy = oy + 2.0f * up.y;
            fx = fmaxf(-p.maxflow, fminf(fx, p.maxflow));
            fy = fmaxf(-p.maxflow, fminf(fy, p.maxflow));
            if (out) {
                st(p.a0_out, p.a0_out_pitch, y, g.x, a0);
                st(p.dflow_out, p.dflow_out_pitch, y, g.x, make_float2(ox, oy));
            }
            ox = fx;
            oy = fy;
        } else {
            const float Yt = ld(p.a0_in, p.a0_in_pitch, yc, g.xc) - a0;
            solve2x2(gx, gy, Yt, s.fu[r], s.fv[r], p.gamma, ox, oy);
            ox = fmaxf(-p.maxflow, fminf(ox, p.maxflow));
            oy = fmaxf(-p.maxflow, fminf(oy, p.maxflow));
            ox = not_finite(ox) ? 0.0f : ox;
            oy = not_finite(oy) ? 0.0f : oy;
            if (out) {
                st(p.a0_out, p.a0_out_pitch, y, g.x, a0);
                if (p.flow_upd) st(p.flow_upd, p.flow_upd_pitch, y, g.x, make_float2(ox, oy));
            }
        }
        s.fu[r] = ox;
        s.fv[r] = oy;
    }

    // ------------------------------------------------------------ smoothing
    // flowSmoothX_k / flowSmoothY_k (flowsmoothing_k.cu:21-87), K times.
    for (int it = 0; it < p.K; ++it) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s.fu[r] = taps_x(s.fu[r], WBox{});
            s.fv[r] = taps_x(s.fv[r], WBox{});
        }
        {
            float* my = &xch[slot][wave][0][0];
            my[0 * LANES + lane] = s.fu[0];
            my[1 * LANES + lane] = s.fv[0];
            my[2 * LANES + lane] = s.fu[1];
            my[3 * LANES + lane] = s.fv[1];
            my[4 * LANES + lane] = s.fu[R - 2];
            my[5 * LANES + lane] = s.fv[R - 2];
            my[6 * LANES + lane] = s.fu[R - 1];
            my[7 * LANES + lane] = s.fv[R - 1];
        }
        __syncthreads();
        // rows -2, -1 (above) and R, R+1 (below)
        float a2u = s.fu[0], a2v = s.fv[0], a1u = s.fu[0], a1v = s.fv[0];
        float b1u = s.fu[R - 1], b1v = s.fv[R - 1], b2u = s.fu[R - 1], b2v = s.fv[R - 1];
        if (wave > 0) {
            const float* o = &xch[slot][wave - 1][0][0];
            a2u = o[4 * LANES + lane];
            a2v = o[5 * LANES + lane];
            a1u = o[6 * LANES + lane];
            a1v = o[7 * LANES + lane];
        }
        if (wave < NW - 1) {
            const float* o = &xch[slot][wave + 1][0][0];
            b1u = o[0 * LANES + lane];
            b1v = o[1 * LANES + lane];
            b2u = o[2 * LANES + lane];
            b2v = o[3 * LANES + lane];
        }
        slot ^= 1;
        float m2u = a2u, m2v = a2v, m1u = a1u, m1v = a1v;  // old rows r-2, r-1
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float zu = s.fu[r], zv = s.fv[r];
            const float p1u = r + 1 < R ? s.fu[r + 1] : b1u, p1v = r + 1 < R ? s.fv[r + 1] : b1v;
            const float p2u = r + 2 < R ? s.fu[r + 2] : (r + 2 == R ? b1u : b2u);
            const float p2v = r + 2 < R ? s.fv[r + 2] : (r + 2 == R ? b1v : b2v);
            float su = 0.0f, sv = 0.0f;
            su = fmaf(kBox, m2u, su);
            sv = fmaf(kBox, m2v, sv);
            su = fmaf(kBox, m1u, su);
            sv = fmaf(kBox, m1v, sv);
            su = fmaf(kBox, zu, su);
            sv = fmaf(kBox, zv, sv);
            su = fmaf(kBox, p1u, su);
            sv = fmaf(kBox, p1v, sv);
            su = fmaf(kBox, p2u, su);
            sv = fmaf(kBox, p2v, sv);
            m2u = m1u;
            m2v = m1v;
            m1u = zu;
            m1v = zv;
            s.fu[r] = su;
            s.fv[r] = sv;
        }
    }

    // ------------------------------------------------- write the new state
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (out_x && out_row(r)) st(p.flow_out, p.flow_out_pitch, g.yw + r, g.x, make_float2(s.fu[r], s.fv[r]));
    }
}

#ifndef FF_FUSED_R
#define FF_FUSED_R 8
#endif
#ifndef FF_FUSED_NW
#define FF_FUSED_NW 16
#endif
constexpr int kR = FF_FUSED_R;
constexpr int kNW = FF_FUSED_NW;

int halo_of(const LevelArgs& a) { return (a.prop_iters > 2 ? a.prop_iters : 2) + 2 * a.smooth_iters; }

}  // namespace

bool fused_level_supported(const LevelArgs& a) {
    const int h0 = halo_of(a);
    if (a.border < 1 || a.prop_iters < 1 || a.smooth_iters < 1) return false;
    if (2 * h0 >= LANES || 2 * h0 >= kNW * kR) return false;
    const int H = a.img.h, W = a.img.w;
    if (W < LANES || H < kNW * kR) return false;
    if (!a.img.data || !a.flow_s_in.data || !a.a0_in.data || !a.flow_s_out.data || !a.a0_out.data) return false;
    if (a.delta && (!a.dflow_in.data || !a.dflow_out.data || !a.upper.data || !a.xmap || !a.ymap ||
                    a.upper.h <= 0 || a.upper.w <= 0))
        return false;
    for (const DevImg* d : {&a.flow_s_in, &a.a0_in, &a.flow_s_out, &a.a0_out})
        if (d->h != H || d->w != W) return false;
    if (a.delta)
        for (const DevImg* d : {&a.dflow_in, &a.dflow_out})
            if (d->h != H || d->w != W) return false;
    if (a.flow_upd_out.data && (a.flow_upd_out.h != H || a.flow_upd_out.w != W)) return false;
    const DevImg* all[] = {&a.img, &a.flow_s_in, &a.a0_in, &a.dflow_in, &a.flow_s_out, &a.a0_out, &a.dflow_out,
                           &a.flow_upd_out, &a.upper};
    for (const DevImg* d : all)
        if (d->pitch > 0x7fffffff) return false;
    return true;
}

void launch_fused_level(const LevelArgs& a, hipStream_t s) {
    ProfileScope ps(a.delta ? "fused_delta" : "fused_top", s);
    FusedParams p{};
    p.img = a.img.data;
    p.img_pitch = (int)a.img.pitch;
    p.flow_in = (const float2*)a.flow_s_in.data;
    p.flow_in_pitch = (int)a.flow_s_in.pitch;
    p.a0_in = (const float*)a.a0_in.data;
    p.a0_in_pitch = (int)a.a0_in.pitch;
    p.dflow_in = (const float2*)a.dflow_in.data;
    p.dflow_in_pitch = (int)a.dflow_in.pitch;
    p.flow_out = (float2*)a.flow_s_out.data;
    p.flow_out_pitch = (int)a.flow_s_out.pitch;
    p.a0_out = (float*)a.a0_out.data;
    p.a0_out_pitch = (int)a.a0_out.pitch;
    p.dflow_out = (float2*)a.dflow_out.data;
    p.dflow_out_pitch = (int)a.dflow_out.pitch;
    p.flow_upd = (float2*)a.flow_upd_out.data;
    p.flow_upd_pitch = (int)a.flow_upd_out.pitch;
    p.upper = (const float2*)a.upper.data;
    p.upper_pitch = (int)a.upper.pitch;
    p.xmap = a.xmap;
    p.ymap = a.ymap;
    p.W = a.img.w;
    p.H = a.img.h;
    p.N = a.prop_iters;
    p.K = a.smooth_iters;
    p.border = a.border;
    p.h0 = halo_of(a);
    p.TW = LANES - 2 * p.h0;
    p.TH = kNW * kR - 2 * p.h0;
    p.tiles_x = (p.W + p.TW - 1) / p.TW;
    p.tiles_y = (p.H + p.TH - 1) / p.TH;
    p.dt = a.dt;
    p.gamma = a.gamma;
    p.maxflow = a.maxflow;
    const dim3 grid(p.tiles_x * p.tiles_y), block(kNW * LANES);
    if (a.delta) {
        if (a.img_u8) hipLaunchKernelGGL((k_fused_level<true, true, kR, kNW>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((k_fused_level<true, false, kR, kNW>), grid, block, 0, s, p);
    } else {
        if (a.img_u8) hipLaunchKernelGGL((k_fused_level<false, true, kR, kNW>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((k_fused_level<false, false, kR, kNW>), grid, block, 0, s, p);
    }
}

}  // namespace ffamd
