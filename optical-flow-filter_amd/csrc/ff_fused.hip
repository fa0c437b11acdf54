// ff_fused.hip -- the fused frame kernel for gfx950.
//
// One launch advances every level of a frame. Each level is a job; its
// output tiles are workgroups of the launch. A tile runs the whole per-level
// recurrence of the reference (FlowFilter::compute, src/gpu/flowfilter.cu:
// 125-153, and the DeltaFlowFilter update path, :429-503): N upwind
// propagation iterations of the previous smoothed flow (with the old
// brightness and delta flow as payloads on delta levels), the image model,
// the 2x2 update and K box smoothing iterations. State is read once and
// written once per frame; everything in between stays in registers.
//
// Jobs and their order. The top level of a pyramid comes first; the level
// below it reads the top level's unsmoothed flow of the same frame
// (flowfilter.cu:663-672), so its tiles wait, after their propagation
// phase and before the update, until every tile of the top level has
// signalled completion (a counter in device memory). Blocks are dispatched
// in order on each XCD and producers never wait, so every producer a
// consumer waits for is running or finished; the wait is also bounded by a
// spin limit. Lower levels read the previous frame's flow of the level
// above (the reference's ascending update order) and do not wait.
//
// Tile geometry: a workgroup is nw waves stacked vertically (nw is a launch
// parameter, picked by timing the candidates); lane l of every wave owns
// buffer column l (64 columns), and each thread owns R consecutive rows of
// its wave's strip (R = 8, or 4 on a top level that the level below waits
// for), so the buffer is 64 x (nw*R) pixels. The output
// tile is the buffer minus a halo h0 = max(N, 2 + e) + 2K on every side (the
// distance an error at the buffer edge travels inward: one pixel per
// propagation pass, two per smoothing pass, two for the image model's x
// taps, plus e = 1 when the image is reduced while loading, because then
// the buffer's first column also lacks its left source pixel).
//  - X neighbours come from DPP wave shifts; Y neighbours are registers of
//    the same thread, and the rows above and below a wave's strip are
//    exchanged through LDS once per Y pass.
//  - On levels at least as large as the buffer every buffer lies inside the
//    image: buffers at the image edges are clamped against the edge and
//    their output range extends to it. At an
//    image edge the missing neighbour must be the clamp-to-edge read of the
//    reference: the first/last wave use their own first/last row, and the
//    x taps of the image model and smoothing use DPP shifts in which the
//    edge lane keeps its own value. The propagation passes use zero-filling
//    shifts (a single instruction the compiler folds into the arithmetic):
//    with border >= 1 the pixels next to an image edge keep their value, so
//    the filled lane only affects pixels inside the halo.
//  - All global accesses are buffer instructions with 32-bit offsets; stores
//    of pixels outside the tile's output range use an offset beyond the
//    buffer's size, which the hardware discards (no branches).
//  - Image rows, the upper-level flow and the old brightness are loaded
//    before they are needed so their latency overlaps computation.
//  - A level narrower or shorter than the buffer is a single tile across or
//    down; the buffer then extends past the image's last column or row,
//    whose values the loads replicate (see level_tile).
//  The fused path needs border >= 1.
//
// Arithmetic per pixel is the reference's (same operations, order and
// rounding points as the stage kernels and the CPU oracle); results are
// bitwise identical to the stage-by-stage path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int LANES = 64;
constexpr int OOB = (int)0x80000000;  // buffer offset beyond any image: access dropped
constexpr int kSpinLimit = 1 << 23;   // polls before a wait gives up: about 2 s at s_sleep 8

// One level of one frame, as the kernel reads it. Blocks first_block ..
// first_block + tiles_x * tiles_y - 1 of the launch are its tiles.
struct LevelJob {
    const void* img;
    const float2* flow_in;
    const float* a0_in;
    const float2* dflow_in;
    float2* flow_out;
    float* a0_out;
    float2* dflow_out;
    float2* flow_upd;  // may be null on a top level
    const float2* upper;
    const int* xmap;
    const int* ymap;
    // row pitches in bytes
    int img_pitch, flow_in_pitch, a0_in_pitch, dflow_in_pitch, flow_out_pitch, a0_out_pitch, dflow_out_pitch,
        flow_upd_pitch, upper_pitch;
    int W, H, Hu, Hsrc;  // level size, upper level height, image height
    int N, K, border, h0, TW, TH, tiles_x, tiles_y;
    int delta, img_mode, first_block;
    // completion counters: wait until counter wait_idx has counted
    // wait_target more tiles in this launch than before it; signal counter
    // signal_idx. frame_slot >= 0: the level's image is a filter's input
    // frame, which a launch may pass in place as its frame frame_slot (read
    // where the caller keeps it).
    int wait_idx, wait_target, frame_slot, signal_idx;
    float dt, gamma, maxflow;
    int pad_;
};

// ------------------------------------------------------------ buffer access

typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ Rsrc rsrc(const void* p, int bytes) {
    // word 3 of the buffer descriptor for gfx9-family raw (untyped) buffers
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
__device__ __forceinline__ float ld1(Rsrc r, int off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ float2 ld2(Rsrc r, int off) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ unsigned ldb(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0); }
__device__ __forceinline__ unsigned ldh(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0); }
// Cache policy sc1: stores write through this XCD's L2 and such loads do
// not hit a stale line in its caches (device-coherent, as for agent-scope
// atomics).
constexpr int kCoherent = 16;
// Every store writes through. A frame's state is written once and read by
// the next launch only, so dirty lines kept in L2 save nothing, and they
// cost time: 55.8 us per 2048x1088 two-level frame with write-back stores,
// 52.8 us with write-through ones (streaming stores and loads, policy nt,
// were slower than either). The write-through also publishes the top
// level's unsmoothed flow to the tiles of the level below, which read it
// within the same launch with ld2_coherent.
__device__ __forceinline__ void st1(Rsrc r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, kCoherent);
}
__device__ __forceinline__ void st2(Rsrc r, int off, float a, float b) {
    const u32x2 v = {__float_as_uint(a), __float_as_uint(b)};
    __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, kCoherent);
}
__device__ __forceinline__ float2 ld2_coherent(Rsrc r, int off) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kCoherent);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

// ---------------------------------------------------------- lane shifts
// Lane i receives lane i-1 (shr) or lane i+1 (shl). The zero-filling forms
// give 0 to the lane without a source; the clamping forms let it keep its
// own value (the clamp-to-edge read at an image edge).
__device__ __forceinline__ float zshr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float zshl(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float cshr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float cshl(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
// Lane i receives lane i-1's v; lane 0 receives `own`. (A select between a
// shifted value and a lane's own value must not be written as a condition:
// the compiler may then run the shift with some lanes disabled, and a lane
// reading a disabled lane receives 0.)
__device__ __forceinline__ unsigned shr_or(unsigned v, unsigned own) {
    return (unsigned)__builtin_amdgcn_update_dpp((int)own, (int)v, 0x138, 0xf, 0xf, false);
}
template <bool CLAMP>
__device__ __forceinline__ float shr(float v) {
    if constexpr (CLAMP) return cshr(v);
    else return zshr(v);
}
template <bool CLAMP>
__device__ __forceinline__ float shl(float v) {
    if constexpr (CLAMP) return cshl(v);
    else return zshl(v);
}

// 5-tap correlation along x in the order of the reference loops (c = -2..2):
// acc = 0; acc = fma(w[c], v(x+c), acc).
template <bool CLAMP, class Wf>
__device__ __forceinline__ float taps_x(float v, Wf w) {
    const float m1 = shr<CLAMP>(v), p1 = shl<CLAMP>(v);
    float a = 0.0f;
    a = fmaf(w(0), shr<CLAMP>(m1), a);
    a = fmaf(w(1), m1, a);
    a = fmaf(w(2), v, a);
    a = fmaf(w(3), p1, a);
    a = fmaf(w(4), shl<CLAMP>(p1), a);
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

// Map a job's tile index so that the tiles of one XCD (blocks are dealt
// round-robin over the 8 XCDs) form a contiguous band of the image.
__device__ __forceinline__ int xcd_tile(int b, int nt) {
    const int q = nt / 8;
    if (b >= 8 * q) return b;
    return (b % 8) * q + b / 8;
}

// ------------------------------------------------- completion counters

// Wait until the producer job's tiles have all signalled. The counters
// only grow: a launch with sequence number `epoch` (1, 2, ... since the host
// last cleared them) waits for epoch * wait_target, compared modulo 2^32.
// Nothing is reset in the kernel, so the wait costs one relaxed load when the
// producer is done (a returning atomic here would put a device-scope round
// trip on every consumer tile's critical path). The data the producers
// published (the unsmoothed flow) is read with device-coherent loads after
// the wait, so no cache invalidation is needed; the workgroup barrier
// orders every wave's loads after the polling thread saw the count.
// The wait gives up after `spin_limit` polls and counts a timeout (the host
// reports it as an error); `bias` is added to the target (0, or a count no
// launch reaches, which a test uses to make the wait give up).
__device__ __forceinline__ void wait_upper(const LevelJob& p, unsigned* counters, int timeout_idx, unsigned epoch,
                                           int spin_limit, unsigned bias) {
    if (threadIdx.x == 0) {
        const unsigned* done = counters + 2 * p.wait_idx;
        const unsigned target = epoch * (unsigned)p.wait_target + bias;
        int spins = 0;
        while ((int)(__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            __builtin_amdgcn_s_sleep(8);
            if (++spins > spin_limit) {
                __hip_atomic_fetch_add(counters + timeout_idx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

// Count this tile as done once its device-coherent stores have completed:
// every wave waits for its outstanding stores (vmcnt 0; the immediate
// leaves the other counters unwaited), the barrier joins the waves, and one
// thread increments the counter.
__device__ __forceinline__ void signal_done(const LevelJob& p, unsigned* counters) {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_fetch_add(counters + 2 * p.signal_idx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------ tile body

// Per-thread view of the tile.
struct Geo {
    int lane, wave, x, yw;
    int xl;        // column read: x, or the last column for x beyond the image
    int xlast;     // lane of the image's last column (>= LANES - 1: not in the buffer)
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
// propagation_k.cu:16-72, 134-232). LAST: the final iteration of a delta
// level, whose propagated flow is read only as the Y pass's speed (the
// delta update consumes the payloads): the u component is not advanced.
template <bool DELTA, bool BORDER, int R, bool LAST = false>
__device__ __forceinline__ void prop_x(State<DELTA, R>& s, const Geo& g, const LevelJob& p) {
    constexpr bool kSkipU = DELTA && LAST;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float zu = s.fu[r], zv = s.fv[r];
        const float mu = zshr(zu), pu = zshl(zu);
        const float Ud = (fabsf(pu) - fabsf(mu)) > 0.0f ? pu : mu;
        const float nd = -(p.dt * Ud);
        const bool back = Ud >= 0.0f;
        float ou = kSkipU ? zu : upwind(mu, zu, pu, back, nd);
        float ov = upwind(zshr(zv), zv, zshl(zv), back, nd);
        float os = 0.f, odu = 0.f, odv = 0.f;
        if constexpr (DELTA) {
            os = upwind(zshr(s.sa[r]), s.sa[r], zshl(s.sa[r]), back, nd);
            odu = upwind(zshr(s.du[r]), s.du[r], zshl(s.du[r]), back, nd);
            odv = upwind(zshr(s.dv[r]), s.dv[r], zshl(s.dv[r]), back, nd);
        }
        if constexpr (BORDER) {
            const int y = g.yw + r;
            const bool in = g.xin && y >= p.border && y < p.H - p.border;
            if constexpr (!kSkipU) ou = in ? ou : zu;
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
// of row r-1. LAST: the final pass of a delta level, whose propagated flow
// is dead (only the payloads feed the delta update): the flow is not
// advanced, only read for the speed Vd.
template <bool DELTA, bool BORDER, int R, bool LAST = false>
__device__ __forceinline__ void prop_y(State<DELTA, R>& s, const Geo& g, const LevelJob& p, const Edge5& a,
                                       const Edge5& b) {
    constexpr bool kSkipF = DELTA && LAST;
    float mu = a.u, mv = a.v, ms = a.s, mdu = a.du, mdv = a.dv;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float zu = s.fu[r], zv = s.fv[r];
        const float pu = r + 1 < R ? s.fu[r + 1] : b.u;
        const float pv = r + 1 < R ? s.fv[r + 1] : b.v;
        const float Vd = (fabsf(pv) - fabsf(mv)) > 0.0f ? pv : mv;
        const float nd = -(p.dt * Vd);
        const bool back = Vd >= 0.0f;
        float ou = kSkipF ? zu : upwind(mu, zu, pu, back, nd);
        float ov = kSkipF ? zv : upwind(mv, zv, pv, back, nd);
        mu = zu;
        mv = zv;
        bool in = true;
        if constexpr (BORDER) {
            const int y = g.yw + r;
            in = g.xin && y >= p.border && y < p.H - p.border;
            if constexpr (!kSkipF) {
                ou = in ? ou : zu;
                ov = in ? ov : zv;
            }
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

// Image model (imagePrefilter_k / imageModel_k, imagemodel_k.cu:31-133) and
// update (flowUpdate_k, update_k.cu:15-79, or deltaFlowUpdate_k, :82-173)
// of every row of the thread. im[j] is image row yw-2+j (clamped reads),
// aold[r] the old brightness (top level) and up[r] the upper-level flow
// (delta level) of row r. Writes the new brightness (and delta flow) of the
// tile's output pixels; leaves the updated flow in s.fu/s.fv.
template <bool DELTA, bool CLAMP, int R>
__device__ __forceinline__ void model_update(State<DELTA, R>& s, const float (&im)[R + 4], const float (&aold)[R],
                                             const float2 (&up)[R], const LevelJob& p, const Rsrc& ra0,
                                             const Rsrc& rdf, const Rsrc& rfu, const int (&offa)[R],
                                             const int (&offf)[R]) {
    float sx[R + 4];
#pragma unroll
    for (int j = 0; j < R + 4; ++j) sx[j] = taps_x<CLAMP>(im[j], WSmooth{});
#pragma unroll
    for (int r = 0; r < R; ++r) {
        float sy = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) sy = fmaf(smooth_w(k), im[r + k], sy);
        const float a0 = taps_x<CLAMP>(sy, WSmooth{});
        const float gx = taps_x<CLAMP>(sy, WDiff{});
        float gy = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) gy = fmaf(diff_w(k), sx[r + k], gy);
        float ox, oy;
        if constexpr (DELTA) {
            const float Yt = s.sa[r] - a0;
            solve2x2(gx, gy, Yt, s.du[r], s.dv[r], p.gamma, ox, oy);
            ox = not_finite(ox) ? 0.0f : ox;
            oy = not_finite(oy) ? 0.0f : oy;
            const float half = 0.5f * p.maxflow;
            ox = fmaxf(-half, fminf(ox, half));
            oy = fmaxf(-half, fminf(oy, half));
            float fx = ox + 2.0f * up[r].x, fy = oy + 2.0f * up[r].y;
            fx = fmaxf(-p.maxflow, fminf(fx, p.maxflow));
            fy = fmaxf(-p.maxflow, fminf(fy, p.maxflow));
            st1(ra0, offa[r], a0);
            st2(rdf, offf[r], ox, oy);
            ox = fx;
            oy = fy;
        } else {
            const float Yt = aold[r] - a0;
            solve2x2(gx, gy, Yt, s.fu[r], s.fv[r], p.gamma, ox, oy);
            ox = fmaxf(-p.maxflow, fminf(ox, p.maxflow));
            oy = fmaxf(-p.maxflow, fminf(oy, p.maxflow));
            ox = not_finite(ox) ? 0.0f : ox;
            oy = not_finite(oy) ? 0.0f : oy;
            st1(ra0, offa[r], a0);
            if (p.flow_upd) st2(rfu, offf[r], ox, oy);
        }
        s.fu[r] = ox;
        s.fv[r] = oy;
    }
}

// Lanes beyond the image's last column (lane xlast) take its values.
template <int R>
__device__ __forceinline__ void copy_last_column(float (&fu)[R], float (&fv)[R], int lane, int xlast) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float eu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fu[r]), xlast));
        const float ev = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fv[r]), xlast));
        fu[r] = lane > xlast ? eu : fu[r];
        fv[r] = lane > xlast ? ev : fv[r];
    }
}

// X pass of the box smoother (flowSmoothX_k, flowsmoothing_k.cu:21-53).
template <bool CLAMP, int R>
__device__ __forceinline__ void smooth_x(float (&fu)[R], float (&fv)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        fu[r] = taps_x<CLAMP>(fu[r], WBox{});
        fv[r] = taps_x<CLAMP>(fv[r], WBox{});
    }
}

// Y pass of the box smoother (flowSmoothY_k, flowsmoothing_k.cu:55-87) over
// the thread's rows; a2/a1 are the rows above, b1/b2 the rows below. With
// YCLAMP, rows after rlast (outside the image) are replaced by row rlast, the
// clamped read; rows after rlast themselves are not needed.
template <bool YCLAMP, int R>
__device__ __forceinline__ void smooth_y(float (&fu)[R], float (&fv)[R], float a2u, float a2v, float a1u, float a1v,
                                         float b1u, float b1v, float b2u, float b2v, int rlast) {
    float m2u = a2u, m2v = a2v, m1u = a1u, m1v = a1v;  // old rows r-2, r-1
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float zu = fu[r], zv = fv[r];
        float p1u = r + 1 < R ? fu[r + 1] : b1u, p1v = r + 1 < R ? fv[r + 1] : b1v;
        float p2u = r + 2 < R ? fu[r + 2] : (r + 2 == R ? b1u : b2u);
        float p2v = r + 2 < R ? fv[r + 2] : (r + 2 == R ? b1v : b2v);
        if constexpr (YCLAMP) {
            p2u = r + 2 <= rlast ? p2u : (r + 1 <= rlast ? p1u : zu);
            p2v = r + 2 <= rlast ? p2v : (r + 1 <= rlast ? p1v : zv);
            p1u = r + 1 <= rlast ? p1u : zu;
            p1v = r + 1 <= rlast ? p1v : zv;
        }
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
        fu[r] = su;
        fv[r] = sv;
    }
}

// Image rows yw-2 .. yw+R+1 (clamped to the level) at the thread's column.
// load_image issues the loads; finish_image turns them into normalized
// values once they are needed, so the latency of the F32 and U8 loads
// overlaps the propagation phase.
//  IMG_F32:     the float texel.
//  IMG_U8:      the u8 texel.
//  IMG_U8_DOWN: the texel of the level one pyramid step below (twice the
//               size): imageDownX_uint8_k then imageDownY_uint8_k
//               (pyramid_k.cu:14-78) evaluated for this texel. Row yy needs
//               source rows 2yy-1 (clamped), 2yy, 2yy+1; each of those the
//               source columns 2x-1 (clamped), 2x, 2x+1, read as one 16-bit
//               load of columns 2x, 2x+1 and the high byte of the left
//               lane's load. With W = Wsrc/2 and H = Hsrc/2 (integer
//               halving) the columns 2x+1 and rows 2yy+1 are always inside
//               the source image. The reduction runs right after the loads
//               (three source rows per texel would not fit in registers
//               through the propagation phase).

// One pyramid step along x of a source row: columns 2x-1 (clamped), 2x, 2x+1.
// Column 2x-1 is the high byte of the left lane's pair; lane 0 takes its own
// low byte instead, which is the clamped read when x = 0 (for a buffer that
// does not start at the image edge lane 0 lies in the halo).
__device__ __forceinline__ unsigned down_x(unsigned pair) {
    const unsigned lo = pair & 0xffu, hi = pair >> 8;
    const unsigned left = (shr_or(pair, pair << 8) >> 8) & 0xffu;
    return down_u8(left, lo, hi);
}

// Registers holding the image rows between the loads and their use: the
// float texels, or the u8 texels packed four to a register (the byte
// conversion instructions read them in place).
template <int MODE, int R>
struct ImageRegs {
    static constexpr int kN = MODE == IMG_F32 ? R + 4 : (R + 4 + 3) / 4;
    unsigned v[kN];
};

// The level's source image: the job's, or the frame passed to the launch.
struct ImgSrc {
    const void* data;
    int pitch;
};

template <int MODE, int R>
__device__ __forceinline__ void load_image(ImageRegs<MODE, R>& ir, const LevelJob& p, const Geo& g, ImgSrc src) {
    const Rsrc rimg = rsrc(src.data, src.pitch * p.Hsrc);
    const int ipitch = src.pitch;
    if constexpr (MODE == IMG_F32) {
#pragma unroll
        for (int j = 0; j < R + 4; ++j) {
            const int yc = clampi(g.yw - 2 + j, 0, p.H - 1);
            ir.v[j] = __builtin_amdgcn_raw_buffer_load_b32(rimg, yc * ipitch + g.xl * 4, 0, 0);
        }
        return;
    } else {
        unsigned t[R + 4];
        if constexpr (MODE == IMG_U8_DOWN2) {
            // source rows 4yy-3 .. 4yy+3 (level-1 rows 2yy-1 .. 2yy+1, each
            // from its three source rows), one 32-bit load of columns
            // 4x .. 4x+3 per row
            const int H1 = p.Hsrc / 2;
#pragma unroll
            for (int j = 0; j < R + 4; ++j) {
                const int yy = clampi(g.yw - 2 + j, 0, p.H - 1);
                unsigned l1a[3], l1b[3];  // level-1 texels at columns 2x, 2x+1 of rows 2yy-1+k
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int r1 = clampi(2 * yy - 1 + k, 0, H1 - 1);
                    unsigned xa[3], xb[3];
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        const int r0 = max(2 * r1 - 1 + m, 0);
                        const unsigned wd = __builtin_amdgcn_raw_buffer_load_b32(rimg, r0 * ipitch + 4 * g.xl, 0, 0);
                        const unsigned b0 = wd & 0xffu, b1 = (wd >> 8) & 0xffu, b2 = (wd >> 16) & 0xffu, b3 = wd >> 24;
                        const unsigned left = (shr_or(wd, wd << 24) >> 24) & 0xffu;  // column 4x-1 (clamped)
                        xa[m] = down_u8(left, b0, b1);
                        xb[m] = down_u8(b1, b2, b3);
                    }
                    l1a[k] = down_u8(xa[0], xa[1], xa[2]);
                    l1b[k] = down_u8(xb[0], xb[1], xb[2]);
                }
                unsigned yv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    // level-1 column 2x-1 is the left lane's 2x+1 (clamped at x = 0)
                    const unsigned lm = shr_or(l1b[k], l1a[k]);
                    yv[k] = down_u8(lm, l1a[k], l1b[k]);
                }
                t[j] = down_u8(yv[0], yv[1], yv[2]);
            }
        } else if constexpr (MODE == IMG_U8_DOWN) {
            unsigned src[R + 4][3];
#pragma unroll
            for (int j = 0; j < R + 4; ++j) {
                const int r1 = 2 * clampi(g.yw - 2 + j, 0, p.H - 1);
                const int r0 = r1 > 0 ? r1 - 1 : 0;
                src[j][0] = ldh(rimg, r0 * ipitch + 2 * g.xl);
                src[j][1] = ldh(rimg, r1 * ipitch + 2 * g.xl);
                src[j][2] = ldh(rimg, (r1 + 1) * ipitch + 2 * g.xl);
            }
#pragma unroll
            for (int j = 0; j < R + 4; ++j) t[j] = down_u8(down_x(src[j][0]), down_x(src[j][1]), down_x(src[j][2]));
        } else {
#pragma unroll
            for (int j = 0; j < R + 4; ++j) {
                const int yc = clampi(g.yw - 2 + j, 0, p.H - 1);
                t[j] = ldb(rimg, yc * ipitch + g.xl);
            }
        }
        if constexpr (MODE == IMG_U8_DOWN || MODE == IMG_U8_DOWN2) {
            // lanes beyond the last column: copies of it (their own reduction
            // would combine a clamped centre with an unclamped left column)
            if (g.xlast < LANES - 1) {
#pragma unroll
                for (int j = 0; j < R + 4; ++j) {
                    const unsigned e = (unsigned)__builtin_amdgcn_readlane((int)t[j], g.xlast);
                    t[j] = g.lane > g.xlast ? e : t[j];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < ImageRegs<MODE, R>::kN; ++k) {
            unsigned w = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * k + i < R + 4) w |= t[4 * k + i] << (8 * i);
            // pack here: left to the compiler, the bytes (or, reduced on
            // load, three source values per row) would stay in separate
            // registers through the propagation phase
            asm volatile("" : "+v"(w));
            ir.v[k] = w;
        }
    }
}

template <int MODE, int R>
__device__ __forceinline__ void finish_image(const ImageRegs<MODE, R>& ir, float (&im)[R + 4]) {
#pragma unroll
    for (int j = 0; j < R + 4; ++j) {
        if constexpr (MODE == IMG_F32) im[j] = __uint_as_float(ir.v[j]);
        else im[j] = unorm8((ir.v[j / 4] >> (8 * (j % 4))) & 0xffu);
    }
}

// Row exchange buffer of the workgroup: [2][nw][10][LANES] floats (double
// buffered; 10 rows of each wave: its first and last rows of up to five
// fields).
__device__ __forceinline__ float* xrow(float* xch, int nw, int slot, int wave) {
    return xch + ((slot * nw + wave) * 10) * LANES;
}

// Per-launch arguments of the frame kernel besides the job table.
struct WaitArgs {
    int spin_limit;
    unsigned bias;
};
struct FrameArgs {
    const unsigned char* ptr[kMaxLaunchFrames];
    int pitch[kMaxLaunchFrames];
};

// INNER: a tile whose every pixel lies inside the border, away from the
// image edges (every wave inside, no edge lanes, no clamped rows, no
// column or row beyond the image): one straight-line body with no per-wave
// choice of pass forms, so no register copies where two forms of a pass
// would join; it computes what the general body computes for such a tile.
// ALLOW_INNER: interior tiles take that body (a kernel instantiation of
// its own, since the second body changes the register allocation of the
// first). Other tiles choose per wave.
template <bool DELTA, int MODE, int R, bool ALLOW_INNER, bool INNER = false>
__device__ __forceinline__ void level_tile(const LevelJob& p, int tile, float* xch, int nw, unsigned* counters, unsigned epoch,
                                           int timeout_idx, const WaitArgs& wa, ImgSrc frame) {
    Geo g;
    g.lane = threadIdx.x & (LANES - 1);
    g.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = xcd_tile(tile, p.tiles_x * p.tiles_y);
    const int tyi = t / p.tiles_x, txi = t - tyi * p.tiles_x;
    // output range of the tile and the buffer around it, clamped into the image
    const int ox0 = txi * p.TW, ox1 = min(ox0 + p.TW, p.W);
    const int oy0 = tyi * p.TH, oy1 = min(oy0 + p.TH, p.H);
    const int x0 = clampi(ox0 - p.h0, 0, p.W - LANES);
    const int y0 = clampi(oy0 - p.h0, 0, p.H - nw * R);
    if constexpr (ALLOW_INNER && !INNER) {
        // at least one column and row from each image edge, also when the
        // propagation border is 0
        const int bb = max(p.border, 1);
        if (x0 >= bb && x0 + LANES - 1 < p.W - bb && y0 >= bb && y0 + nw * R - 1 < p.H - bb) {
            level_tile<DELTA, MODE, R, true, true>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, frame);
            return;
        }
    }
    g.x = x0 + g.lane;
    g.xl = min(g.x, p.W - 1);
    g.yw = y0 + g.wave * R;  // image row of this thread's first row (wave-uniform)
    const int b = p.border;
    g.xin = g.x >= b && g.x < p.W - b;
    g.wave_in = (x0 >= b) && (x0 + LANES - 1 < p.W - b) && (g.yw >= b) && (g.yw + R - 1 < p.H - b);
    // an image edge at or inside the buffer's left/right edge
    const bool xedge = (x0 == 0) || (x0 + LANES >= p.W);
    // Images narrower than the buffer (a single tile across): the lanes
    // beyond the last column hold copies of it, loaded by clamped reads, so
    // the image model's taps and the propagation (which keeps pixels within
    // the border) see the reference's clamp-to-edge reads; the smoothing
    // passes re-copy the last column before every X pass and clamp their row
    // neighbours at the last row (rlast: the thread's last row inside the
    // image, relative to its first row).
    const int xlast = p.W - 1 - x0;  // lane of the last column
    const bool xinside = xlast < LANES - 1;
    g.xlast = xlast;
    const int rlast = p.H - 1 - g.yw;
    const bool yinside = y0 + nw * R > p.H;
    const int lane = g.lane, wave = g.wave;
    const bool out_x = g.x >= ox0 && g.x < ox1;

    const Rsrc rfi = rsrc(p.flow_in, p.flow_in_pitch * p.H);
    const Rsrc rai = rsrc(p.a0_in, p.a0_in_pitch * p.H);
    const Rsrc rfo = rsrc(p.flow_out, p.flow_out_pitch * p.H);
    const Rsrc rao = rsrc(p.a0_out, p.a0_out_pitch * p.H);

    State<DELTA, R> s;

    // ------------------------------------- load the state and the image rows
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = min(g.yw + r, p.H - 1);
        const float2 f = ld2(rfi, y * p.flow_in_pitch + g.xl * 8);
        s.fu[r] = f.x;
        s.fv[r] = f.y;
        if constexpr (DELTA) {
            s.sa[r] = ld1(rai, y * p.a0_in_pitch + g.xl * 4);
            const float2 d = ld2(rsrc(p.dflow_in, p.dflow_in_pitch * p.H), y * p.dflow_in_pitch + g.xl * 8);
            s.du[r] = d.x;
            s.dv[r] = d.y;
        }
    }
    ImageRegs<MODE, R> ir;
    load_image<MODE, R>(ir, p, g, frame.data ? frame : ImgSrc{p.img, p.img_pitch});

    int slot = 0;

    // ---------------------------------------------------------- propagation
    // N iterations of an X and a Y pass; the last one separately, since on
    // a delta level part of its flow output is dead
    auto iteration = [&](auto last) {
        constexpr bool LAST = decltype(last)::value;
        if (INNER || g.wave_in) prop_x<DELTA, false, R, LAST>(s, g, p);
        else prop_x<DELTA, true, R, LAST>(s, g, p);
        {
            float* my = xrow(xch, nw, slot, wave);
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
            const float* o = xrow(xch, nw, slot, wave - 1);
            ea.u = o[5 * LANES + lane];
            ea.v = o[6 * LANES + lane];
            if constexpr (DELTA) {
                ea.s = o[7 * LANES + lane];
                ea.du = o[8 * LANES + lane];
                ea.dv = o[9 * LANES + lane];
            }
        }
        if (wave < nw - 1) {
            const float* o = xrow(xch, nw, slot, wave + 1);
            eb.u = o[0 * LANES + lane];
            eb.v = o[1 * LANES + lane];
            if constexpr (DELTA) {
                eb.s = o[2 * LANES + lane];
                eb.du = o[3 * LANES + lane];
                eb.dv = o[4 * LANES + lane];
            }
        }
        slot ^= 1;
        if (INNER || g.wave_in) prop_y<DELTA, false, R, LAST>(s, g, p, ea, eb);
        else prop_y<DELTA, true, R, LAST>(s, g, p, ea, eb);
    };
    for (int it = 0; it + 1 < p.N; ++it) iteration(std::false_type{});
    iteration(std::true_type{});

    // ------------------------------------------------ image model + update
    if (DELTA && p.wait_idx >= 0) wait_upper(p, counters, timeout_idx, epoch, wa.spin_limit, wa.bias);
    float aold[R];  // unused (and removed) on delta levels
    float2 up[R];
    int offa[R], offf[R];
    const Rsrc rdf = rsrc(p.dflow_out, p.dflow_out_pitch * p.H);
    const Rsrc rfu = rsrc(p.flow_upd, p.flow_upd_pitch * p.H);
    const int xmx = DELTA ? p.xmap[g.xl] : 0;
    const Rsrc rup = rsrc(p.upper, p.upper_pitch * p.Hu);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = g.yw + r;
        const bool out = out_x && y >= oy0 && y < oy1;
        const int yl = min(y, p.H - 1);
        offa[r] = out ? y * p.a0_out_pitch + g.x * 4 : OOB;
        if constexpr (DELTA) {
            offf[r] = out ? y * p.dflow_out_pitch + g.x * 8 : OOB;
            up[r] = ld2_coherent(rup, p.ymap[yl] * p.upper_pitch + xmx * 8);
            aold[r] = 0.f;
        } else {
            offf[r] = out ? y * p.flow_upd_pitch + g.x * 8 : OOB;
            aold[r] = ld1(rai, yl * p.a0_in_pitch + g.xl * 4);
            up[r] = make_float2(0.f, 0.f);
        }
    }
    float im[R + 4];
    finish_image<MODE, R>(ir, im);
    if (!INNER && xedge) model_update<DELTA, true, R>(s, im, aold, up, p, rao, rdf, rfu, offa, offf);
    else model_update<DELTA, false, R>(s, im, aold, up, p, rao, rdf, rfu, offa, offf);

    // ------------------------------------------------------------ smoothing
    // flowSmoothX_k / flowSmoothY_k (flowsmoothing_k.cu:21-87), K times.
    for (int it = 0; it < p.K; ++it) {
        if (!INNER && xinside) copy_last_column<R>(s.fu, s.fv, g.lane, xlast);
        if (!INNER && xedge) smooth_x<true, R>(s.fu, s.fv);
        else smooth_x<false, R>(s.fu, s.fv);
        {
            float* my = xrow(xch, nw, slot, wave);
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
            const float* o = xrow(xch, nw, slot, wave - 1);
            a2u = o[4 * LANES + lane];
            a2v = o[5 * LANES + lane];
            a1u = o[6 * LANES + lane];
            a1v = o[7 * LANES + lane];
        }
        if (wave < nw - 1) {
            const float* o = xrow(xch, nw, slot, wave + 1);
            b1u = o[0 * LANES + lane];
            b1v = o[1 * LANES + lane];
            b2u = o[2 * LANES + lane];
            b2v = o[3 * LANES + lane];
        }
        slot ^= 1;
        if (!INNER && yinside) smooth_y<true, R>(s.fu, s.fv, a2u, a2v, a1u, a1v, b1u, b1v, b2u, b2v, rlast);
        else smooth_y<false, R>(s.fu, s.fv, a2u, a2v, a1u, a1v, b1u, b1v, b2u, b2v, rlast);
    }

    // ------------------------------------------------- write the new state
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = g.yw + r;
        const bool out = out_x && y >= oy0 && y < oy1;
        st2(rfo, out ? y * p.flow_out_pitch + g.x * 8 : OOB, s.fu[r], s.fv[r]);
    }
    if (!DELTA && p.signal_idx >= 0) signal_done(p, counters);
}

// Blocks are grouped by tile body: top levels with float, u8, reduced u8
// and twice-reduced u8 images, then delta levels with float, u8 and reduced
// u8 images; ends[i] is the first block after group i. (Selecting the body
// from the block index rather than from a job field also keeps the
// compiler's register allocation of each body as tight as in a kernel of
// its own.)
struct GroupEnds {
    int e[6];
};

// Rows per thread of a tile body: RLOW, except for top levels reduced on
// load (the top level of a pyramid, which the level below waits for), which
// use RTOP; both are parameters of the kernel, chosen per launch by the
// tuner among kShapes. RTOP 4: the top level's tiles are more and shorter
// and finish sooner (a pyramid launched alone is faster); 8: they carry less
// halo (a 64-stream batch is 1.4 % faster). RLOW 4 with 16 waves covers
// the same 64-row tile as RLOW 8 with 8 waves, in half the registers per
// thread: that kernel is held to 64 VGPRs, so two 16-wave workgroups (32
// waves) fit a CU instead of two 8-wave ones, for about 8 % more VALU
// instructions per pixel (each wave exchanges its edge rows for half as
// many rows).
constexpr int kRows = 8;
constexpr int kTopRowChoices[] = {4, 8};
struct TileShape {
    int rtop, rlow;
    bool inner;  // interior tiles in a body of their own (k_fused_frame's INNER)
};
constexpr TileShape kShapes[] = {{4, 8, false}, {8, 8, false}, {4, 4, false}, {4, 4, true}};

template <int RTOP, int RLOW = kRows, bool INNER = false>
__global__ void __launch_bounds__(1024, RLOW == 4 ? 8 : 1)
    k_fused_frame(const LevelJob* __restrict__ jobs, const int* __restrict__ block_job, unsigned* counters,
                  unsigned epoch, int timeout_idx, GroupEnds ends, int nw, WaitArgs wa, FrameArgs frames) {
    // row exchange between the waves of the workgroup (2 * nw * 10 * LANES floats)
    extern __shared__ float xch[];
    const LevelJob& p = jobs[block_job[blockIdx.x]];
    const int tile = (int)blockIdx.x - p.first_block;
    const unsigned b = blockIdx.x;
    // the job's input frame passed in place, if any (frame_slot is
    // uniform, so is the selection)
    ImgSrc fr{nullptr, 0};
    if (p.frame_slot >= 0) fr = ImgSrc{frames.ptr[p.frame_slot], frames.pitch[p.frame_slot]};
    if (b < (unsigned)ends.e[0]) level_tile<false, IMG_F32, RLOW, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else if (b < (unsigned)ends.e[1]) level_tile<false, IMG_U8, RLOW, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else if (b < (unsigned)ends.e[2]) level_tile<false, IMG_U8_DOWN, RTOP, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else if (b < (unsigned)ends.e[3]) level_tile<false, IMG_U8_DOWN2, RTOP, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else if (b < (unsigned)ends.e[4]) level_tile<true, IMG_F32, RLOW, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else if (b < (unsigned)ends.e[5]) level_tile<true, IMG_U8, RLOW, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
    else level_tile<true, IMG_U8_DOWN, RLOW, INNER>(p, tile, xch, nw, counters, epoch, timeout_idx, wa, fr);
}

__global__ void k_clear_counters(unsigned* c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0u;
}

constexpr int kMaxWaves = 16;  // waves per workgroup, at most
// Waves per workgroup the autotuner tries: 16 and 15 (one workgroup per CU),
// 8 (two), 5 (three) and 4 (four); the register budget allows 16 waves per
// CU.
constexpr int kWaveChoices[] = {16, 15, 8, 5, 4};

// Whether a job runs a top-level body reduced on load (rows per thread
// RTOP).
bool top_rows_body(const LevelArgs& a) { return !a.delta && (a.img_mode == IMG_U8_DOWN || a.img_mode == IMG_U8_DOWN2); }

// Rows per thread of a job's tile body (see k_fused_frame).
int rows_per_thread(const LevelArgs& a, int rtop, int rlow = kRows) { return top_rows_body(a) ? rtop : rlow; }

size_t lds_bytes(int nw) { return (size_t)2 * nw * 10 * LANES * sizeof(float); }

// Level size of a job (the image is larger when reduced on load).
void level_size(const LevelArgs& a, int& H, int& W) {
    H = a.flow_s_in.h;
    W = a.flow_s_in.w;
}

int halo_of(const LevelArgs& a) {
    const int lo = (a.img_mode == IMG_U8_DOWN || a.img_mode == IMG_U8_DOWN2) ? 3 : 2;
    return (a.prop_iters > lo ? a.prop_iters : lo) + 2 * a.smooth_iters;
}

bool fits_i32(const DevImg& d) {
    return (unsigned long long)d.pitch * (unsigned long long)(d.h > 0 ? d.h : 0) < 0x7fffffffULL;
}

// The tiling of a level with buffers of `rows` rows leaves a non-empty
// output tile (or the level fits in one tile).
bool tiles_with(const LevelArgs& a, int rows) {
    int H, W;
    level_size(a, H, W);
    const int h0 = halo_of(a);
    return !((W > LANES && 2 * h0 >= LANES) || (H > rows && 2 * h0 >= rows));
}

}  // namespace

bool fused_level_supported(const LevelArgs& a) {
    if (a.border < 1 || a.prop_iters < 1 || a.smooth_iters < 1) return false;
    int H, W;
    level_size(a, H, W);
    if (H < 1 || W < 1) return false;
    if (!tiles_with(a, kMaxWaves * rows_per_thread(a, kTopRowChoices[0]))) return false;
    if (!a.img.data || !a.flow_s_in.data || !a.a0_in.data || !a.flow_s_out.data || !a.a0_out.data) return false;
    switch (a.img_mode) {
    case IMG_F32:
    case IMG_U8:
        if (a.img.h != H || a.img.w != W) return false;
        break;
    case IMG_U8_DOWN:
        if (a.img.h / 2 != H || a.img.w / 2 != W) return false;
        break;
    case IMG_U8_DOWN2:
        if ((a.img.h / 2) / 2 != H || (a.img.w / 2) / 2 != W || a.delta) return false;
        break;
    default:
        return false;
    }
    if (a.delta && (!a.dflow_in.data || !a.dflow_out.data || !a.upper.data || !a.xmap || !a.ymap ||
                    a.upper.h <= 0 || a.upper.w <= 0))
        return false;
    if (!a.delta && a.wait_on >= 0) return false;
    for (const DevImg* d : {&a.a0_in, &a.flow_s_out, &a.a0_out})
        if (d->h != H || d->w != W) return false;
    if (a.delta)
        for (const DevImg* d : {&a.dflow_in, &a.dflow_out})
            if (d->h != H || d->w != W) return false;
    if (a.flow_upd_out.data && (a.flow_upd_out.h != H || a.flow_upd_out.w != W)) return false;
    // every buffer must be addressable with 32-bit offsets
    for (const DevImg* d : {&a.img, &a.flow_s_in, &a.a0_in, &a.dflow_in, &a.flow_s_out, &a.a0_out, &a.dflow_out,
                            &a.flow_upd_out, &a.upper})
        if (!fits_i32(*d)) return false;
    return true;
}

// ------------------------------------------------------------- FusedFrame

namespace {
constexpr int kGroups = 7;
// Tile body group of a job (the order of the kernel's dispatch).
int body_group(const LevelArgs& a) {
    if (!a.delta) return a.img_mode;  // IMG_F32 .. IMG_U8_DOWN2: groups 0 .. 3
    return 4 + a.img_mode;            // IMG_F32 .. IMG_U8_DOWN: groups 4 .. 6
}

// What the tiling choice depends on: the shape and halo of every level.
std::vector<int> tune_key(const std::vector<LevelArgs>& jobs) {
    std::vector<int> k;
    for (const LevelArgs& a : jobs) {
        int H, W;
        level_size(a, H, W);
        k.insert(k.end(), {H, W, halo_of(a), a.prop_iters, a.smooth_iters, a.delta, a.img_mode, a.wait_on});
    }
    return k;
}
}  // namespace

struct FusedFrame::Impl {
    struct Table {
        std::vector<char> key;  // job records + block map + launch shape, as uploaded
        DeviceArray dev;
        int njobs = 0, blocks = 0, nw = 0;
        GroupEnds ends{};
        int rtop = 0, rlow = kRows;
        bool inner = false;  // launched as k_fused_frame<4, 4, true>
        size_t map_offset = 0;
        unsigned long long token = 0;
        bool waits = false;  // some job waits for another
    };
    Table slot[2];
    int next = 0;
    unsigned long long uploads = 0;
    // Completion counters: one region of `npairs` (done, spare) pairs per
    // table slot (the two slots alternate when a filter's state buffers
    // ping-pong between frames), then the timeout count.
    DeviceArray counters;
    int npairs = 0;
    // per slot: launches counted since its region was last cleared, and the
    // table they count for (0: clear before the next launch)
    unsigned epoch[2] = {0, 0};
    unsigned long long counted[2] = {0, 0};
    // waves per workgroup, top-level rows per thread, other rows per thread;
    // 0: autotune
    int fixed = 0, fixed_rtop = 0, fixed_rlow = 0;
    int fixed_inner = 0;  // 0: tuned, 1: interior tiles in their own body, 2: not
    std::vector<int> tuned_for;
    int tuned = 0, tuned_rtop = 0, tuned_rlow = kRows;
    bool tuned_inner = false;
    std::shared_ptr<WaitWatch> watch = std::make_shared<WaitWatch>();
    int test_spins = 0;  // > 0: the waits give up after this many polls (test hook)

    int table(const std::vector<LevelArgs>& jobs, int nw, int rtop, int rlow, bool inner, hipStream_t s);
    unsigned* region(int si) const { return (unsigned*)counters.data() + (size_t)si * 2 * npairs; }
    int timeout_word() const { return 4 * npairs; }
    void clear_counters(int si, hipStream_t s) const;
    void run(int t, hipStream_t s, const FrameRef* frames = nullptr, int nframes = 0);
    void tune(const std::vector<LevelArgs>& jobs, hipStream_t s);
};

FusedFrame::FusedFrame() : impl_(new Impl) {}
FusedFrame::~FusedFrame() { delete impl_; }
FusedFrame::FusedFrame(FusedFrame&& o) noexcept : impl_(o.impl_) { o.impl_ = new Impl; }
FusedFrame& FusedFrame::operator=(FusedFrame&& o) noexcept {
    std::swap(impl_, o.impl_);
    return *this;
}

// Job table of `jobs` tiled with nw waves per workgroup and rtop rows per
// thread on reduced top levels: found among the two resident tables or
// built and uploaded (after waiting for stream s, the stream of every launch
// of this FusedFrame, since the slot it replaces may be in use).
int FusedFrame::Impl::table(const std::vector<LevelArgs>& jobs, int nw, int rtop, int rlow, bool inner,
                            hipStream_t s) {
    // launch order: grouped by tile body, stable within a group; every top
    // level precedes every delta level, so producers precede consumers
    std::vector<int> order(jobs.size()), pos(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return body_group(jobs[x]) < body_group(jobs[y]); });
    for (size_t i = 0; i < order.size(); ++i) pos[order[i]] = (int)i;
    std::vector<LevelJob> rec(jobs.size());
    int blocks = 0;
    int ends[kGroups - 1] = {0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < jobs.size(); ++i) {
        const LevelArgs& a = jobs[order[i]];
        const int rows = nw * rows_per_thread(a, rtop, rlow);
        if (!fused_level_supported(a) || !tiles_with(a, rows))
            throw logic_error("fused frame: unsupported level parameters");
        LevelJob& p = rec[i];
        std::memset(&p, 0, sizeof(p));
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
        level_size(a, p.H, p.W);
        p.Hu = a.upper.h;
        p.Hsrc = a.img.h;
        p.N = a.prop_iters;
        p.K = a.smooth_iters;
        p.border = a.border;
        p.h0 = halo_of(a);
        // a level no wider (taller) than the buffer is one tile across (down)
        p.TW = p.W <= LANES ? p.W : LANES - 2 * p.h0;
        p.TH = p.H <= rows ? p.H : rows - 2 * p.h0;
        p.tiles_x = (p.W + p.TW - 1) / p.TW;
        p.tiles_y = (p.H + p.TH - 1) / p.TH;
        p.delta = a.delta;
        p.img_mode = a.img_mode;
        p.first_block = blocks;
        p.wait_idx = -1;
        p.signal_idx = -1;
        p.frame_slot = a.frame_slot < kMaxLaunchFrames ? a.frame_slot : -1;
        p.dt = a.dt;
        p.gamma = a.gamma;
        p.maxflow = a.maxflow;
        blocks += p.tiles_x * p.tiles_y;
        for (int g = body_group(a); g < kGroups - 1; ++g) ends[g] = blocks;
    }
    for (size_t i = 0; i < jobs.size(); ++i) {
        const int k = jobs[i].wait_on;
        if (k < 0) continue;
        if (k >= (int)i) throw logic_error("fused frame: a level waits on a later level");
        if (!jobs[i].delta || jobs[k].delta) throw logic_error("fused frame: only delta levels wait on a top level");
        LevelJob& c = rec[pos[i]];
        LevelJob& q = rec[pos[k]];
        c.wait_idx = pos[k];
        c.wait_target = q.tiles_x * q.tiles_y;
        q.signal_idx = pos[k];
    }
    std::vector<int> map(blocks);
    for (size_t i = 0; i < rec.size(); ++i)
        for (int b = 0; b < rec[i].tiles_x * rec[i].tiles_y; ++b) map[rec[i].first_block + b] = (int)i;

    const size_t jobs_bytes = rec.size() * sizeof(LevelJob);
    const size_t map_offset = (jobs_bytes + 255) & ~size_t(255);
    std::vector<char> key(map_offset + map.size() * sizeof(int), 0);
    std::memcpy(key.data(), rec.data(), jobs_bytes);
    std::memcpy(key.data() + map_offset, map.data(), map.size() * sizeof(int));
    // launch shape: part of the table's identity
    for (int g = 0; g < kGroups - 1; ++g) {
        const char* e = reinterpret_cast<const char*>(&ends[g]);
        key.insert(key.end(), e, e + sizeof(int));
    }
    key.insert(key.end(), reinterpret_cast<const char*>(&nw), reinterpret_cast<const char*>(&nw) + sizeof(int));
    key.insert(key.end(), reinterpret_cast<const char*>(&rtop), reinterpret_cast<const char*>(&rtop) + sizeof(int));
    key.insert(key.end(), reinterpret_cast<const char*>(&rlow), reinterpret_cast<const char*>(&rlow) + sizeof(int));
    key.push_back(inner ? 1 : 0);

    for (int i = 0; i < 2; ++i)
        if (slot[i].key == key) return i;

    // a table (or the counters) may be in use by launches still in flight
    // on the stream
    FF_HIP(hipStreamSynchronize(s));
    const int need = (int)jobs.size();
    if (npairs < need) {
        const int n = need < 64 ? 64 : need;
        // the timed-out waits counted so far are kept
        watch->stream = s;
        watch->read();
        std::vector<unsigned> zero(4 * (size_t)n + 2, 0u);
        counters.assign(zero.data(), zero.size() * sizeof(unsigned));
        npairs = n;
        watch->retire((const unsigned*)counters.data() + timeout_word());
        // launches recorded with the old counters must not be replayed
        for (Table& o : slot) o.token = ++uploads;
    }
    const int si = next;
    Table& t = slot[si];
    t.dev.assign(key.data(), key.size());
    t.key = std::move(key);
    t.njobs = (int)rec.size();
    t.blocks = blocks;
    t.nw = nw;
    t.rtop = rtop;
    t.rlow = rlow;
    t.inner = inner;
    t.map_offset = map_offset;
    t.waits = false;
    for (const LevelJob& r : rec) t.waits = t.waits || r.wait_idx >= 0;
    for (int g = 0; g < kGroups - 1; ++g) t.ends.e[g] = ends[g];
    t.token = ++uploads;
    next ^= 1;
    return si;
}

void FusedFrame::Impl::clear_counters(int si, hipStream_t s) const {
    const int n = 2 * npairs;
    hipLaunchKernelGGL(k_clear_counters, dim3((n + 255) / 256), dim3(256), 0, s, region(si), n);
}

void FusedFrame::Impl::run(int ti, hipStream_t s, const FrameRef* frames, int nframes) {
    const int si = ti & 1;
    const Table& t = slot[si];
    if (!t.blocks) throw logic_error("fused frame: launch of an unprepared table");
    static const hipError_t lds_ok4 = hipFuncSetAttribute(
        (const void*)k_fused_frame<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(kMaxWaves));
    static const hipError_t lds_ok8 = hipFuncSetAttribute(
        (const void*)k_fused_frame<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(kMaxWaves));
    static const hipError_t lds_ok44 = hipFuncSetAttribute(
        (const void*)k_fused_frame<4, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(kMaxWaves));
    static const hipError_t lds_ok44i = hipFuncSetAttribute((const void*)k_fused_frame<4, 4, true>,
                                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                                            (int)lds_bytes(kMaxWaves));
    if (lds_ok4 != hipSuccess) throw_hip(lds_ok4, "hipFuncSetAttribute", __FILE__, __LINE__);
    if (lds_ok8 != hipSuccess) throw_hip(lds_ok8, "hipFuncSetAttribute", __FILE__, __LINE__);
    if (lds_ok44 != hipSuccess) throw_hip(lds_ok44, "hipFuncSetAttribute", __FILE__, __LINE__);
    if (lds_ok44i != hipSuccess) throw_hip(lds_ok44i, "hipFuncSetAttribute", __FILE__, __LINE__);
    // Sequence number of this launch for its slot's completion counters. A
    // launch recorded into a graph clears them itself and counts as the
    // first, so every replay is self-contained; direct launches count up
    // from the last clear, which happens whenever the region served another
    // table or a replay (replayed()) in between.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FF_HIP(hipStreamIsCapturing(s, &cs));
    unsigned e;
    if (cs != hipStreamCaptureStatusNone) {
        clear_counters(si, s);
        e = 1;
        counted[si] = 0;
    } else {
        if (counted[si] != t.token || epoch[si] == 0xffffffffu) {
            clear_counters(si, s);
            epoch[si] = 0;
            counted[si] = t.token;
        }
        e = ++epoch[si];
    }
    const WaitArgs wa{test_spins > 0 ? test_spins : kSpinLimit, test_spins > 0 ? 0x40000000u : 0u};
    FrameArgs fa;
    std::memset(&fa, 0, sizeof(fa));
    for (int k = 0; k < nframes; ++k) {
        fa.ptr[k] = (const unsigned char*)frames[k].ptr;
        fa.pitch[k] = (int)frames[k].pitch;
    }
    watch->stream = s;
    if (t.waits) watch->pending = true;
    const WaitingLaunch order(s, t.waits);
    ProfileScope ps("fused_frame", s);
    const char* base = (const char*)t.dev.data();
    // the timeout word, relative to this slot's region
    const int timeout_idx = timeout_word() - si * 2 * npairs;
    if (t.rlow == 4 && t.inner)
        hipLaunchKernelGGL((k_fused_frame<4, 4, true>), dim3(t.blocks), dim3(t.nw * LANES), lds_bytes(t.nw), s,
                           (const LevelJob*)base, (const int*)(base + t.map_offset), region(si), e, timeout_idx, t.ends,
                           t.nw, wa, fa);
    else if (t.rlow == 4)
        hipLaunchKernelGGL((k_fused_frame<4, 4>), dim3(t.blocks), dim3(t.nw * LANES), lds_bytes(t.nw), s,
                           (const LevelJob*)base, (const int*)(base + t.map_offset), region(si), e, timeout_idx, t.ends,
                           t.nw, wa, fa);
    else if (t.rtop == 8)
        hipLaunchKernelGGL(k_fused_frame<8>, dim3(t.blocks), dim3(t.nw * LANES), lds_bytes(t.nw), s,
                           (const LevelJob*)base, (const int*)(base + t.map_offset), region(si), e, timeout_idx, t.ends,
                           t.nw, wa, fa);
    else
        hipLaunchKernelGGL(k_fused_frame<4>, dim3(t.blocks), dim3(t.nw * LANES), lds_bytes(t.nw), s,
                           (const LevelJob*)base, (const int*)(base + t.map_offset), region(si), e, timeout_idx, t.ends,
                           t.nw, wa, fa);
}

// Time the launches of every candidate tiling (waves per workgroup and, when
// a job runs a reduced top-level body, its rows per thread) on the actual
// job table and keep the fastest; a fixed choice is not varied. A launch
// reads the current state and writes the next one without committing it, so
// repeated launches recompute the same output; every tiling gives bitwise
// the same results.
void FusedFrame::Impl::tune(const std::vector<LevelArgs>& jobs, hipStream_t s) {
    bool reduced_top = false;
    for (const LevelArgs& a : jobs) reduced_top = reduced_top || top_rows_body(a);
    int best_nw = 0, best_rtop = 0, best_rlow = kRows;
    bool best_inner = false;
    float best_ms = 0.0f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FF_HIP(hipEventCreate(&e0));
    FF_HIP(hipEventCreate(&e1));
    try {
        for (const TileShape& sh : kShapes) {
            const int rtop = sh.rtop, rlow = sh.rlow;
            if (fixed_rtop ? rtop != fixed_rtop : (!reduced_top && rtop != kTopRowChoices[0])) continue;
            if (fixed_rlow && rlow != fixed_rlow) continue;
            if (fixed_inner && sh.inner != (fixed_inner == 1)) continue;
            for (int nw : kWaveChoices) {
                if (fixed && nw != fixed) continue;
                // four rows per thread only in the tallest workgroups (the
                // 64-row tiles of eight waves of eight rows)
                if (rlow == 4 && nw < 15) continue;
                bool ok = true;
                for (const LevelArgs& a : jobs) ok = ok && tiles_with(a, nw * rows_per_thread(a, rtop, rlow));
                if (!ok) continue;
                const int t = table(jobs, nw, rtop, rlow, sh.inner, s);
                run(t, s);  // warm-up
                constexpr int kReps = 3;
                FF_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < kReps; ++r) run(t, s);
                FF_HIP(hipEventRecord(e1, s));
                FF_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                FF_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (!best_nw || ms < best_ms) {
                    best_nw = nw;
                    best_rtop = rtop;
                    best_rlow = rlow;
                    best_inner = sh.inner;
                    best_ms = ms;
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!best_nw && fixed_inner == 1)
        throw invalid_argument("fused frame: the interior tile body needs four rows per thread and 16 or 15 waves per "
                               "workgroup, which the fixed part of the tiling excludes");
    if (!best_nw) throw logic_error("fused frame: no tiling fits the levels");
    tuned = best_nw;
    tuned_rtop = best_rtop;
    tuned_rlow = best_rlow;
    tuned_inner = best_inner;
}

int FusedFrame::prepare(const std::vector<LevelArgs>& jobs, hipStream_t s) {
    if (jobs.empty()) throw invalid_argument("fused frame: no levels");
    // the rows below the top level default to eight when the tile height
    // and the top rows are both fixed; otherwise the tuner picks them too
    int nw = impl_->fixed, rtop = impl_->fixed_rtop, rlow = impl_->fixed_rlow ? impl_->fixed_rlow : kRows;
    bool inner = impl_->fixed_inner == 1;
    if (rlow == 4 && nw && nw < 15)
        throw invalid_argument("fused frame: four rows per thread need 16 or 15 waves per workgroup, got " +
                               std::to_string(nw));
    if (inner && nw && rtop && rlow != 4)
        throw invalid_argument("fused frame: the interior tile body needs four rows per thread below the top level");
    if (!nw || !rtop) {
        // tuned per set of level shapes and per fixed part of the choice
        std::vector<int> k = tune_key(jobs);
        k.push_back(impl_->fixed);
        k.push_back(impl_->fixed_rtop);
        k.push_back(impl_->fixed_rlow);
        k.push_back(impl_->fixed_inner);
        if (k != impl_->tuned_for || !impl_->tuned) {
            impl_->tune(jobs, s);
            impl_->tuned_for = std::move(k);
        }
        nw = impl_->tuned;
        rtop = impl_->tuned_rtop;
        rlow = impl_->tuned_rlow;
        inner = impl_->tuned_inner;
    }
    return impl_->table(jobs, nw, rtop, rlow, inner, s);
}

void FusedFrame::launch(int table, hipStream_t s, const FrameRef* frames, int nframes) {
    if (nframes < 0 || nframes > kMaxLaunchFrames) throw invalid_argument("fused frame: too many frames in place");
    for (int k = 0; k < nframes; ++k)
        if (frames[k].ptr && (frames[k].pitch > 0x7fffffffu || frames[k].pitch % 4 || (uintptr_t)frames[k].ptr % 4))
            throw invalid_argument("fused frame: a frame in place needs a pitch and address that are multiples of 4 "
                                   "and a pitch below 2^31");
    impl_->run(table, s, frames, nframes);
}

void FusedFrame::replayed() { impl_->counted[0] = impl_->counted[1] = 0; }

int FusedFrame::blocks(int table) const { return impl_->slot[table & 1].blocks; }
unsigned long long FusedFrame::token(int table) const { return impl_->slot[table & 1].token; }
bool FusedFrame::waits(int table) const { return impl_->slot[table & 1].waits; }

void FusedFrame::setWaves(int nw) {
    bool ok = nw == 0;
    for (int c : kWaveChoices) ok = ok || c == nw;
    if (!ok)
        throw invalid_argument("fused frame: waves per workgroup must be 0 (autotune), 16, 15, 8, 5 or 4, got " +
                               std::to_string(nw));
    impl_->fixed = nw;
}

int FusedFrame::waves() const { return impl_->fixed ? impl_->fixed : impl_->tuned; }

void FusedFrame::setTopRows(int rows) {
    bool ok = rows == 0;
    for (int c : kTopRowChoices) ok = ok || c == rows;
    if (!ok)
        throw invalid_argument("fused frame: top-level rows per thread must be 0 (autotune), 4 or 8, got " +
                               std::to_string(rows));
    impl_->fixed_rtop = rows;
}

int FusedFrame::topRows() const { return impl_->fixed_rtop ? impl_->fixed_rtop : impl_->tuned_rtop; }
int FusedFrame::rows() const {
    if (impl_->fixed_rlow) return impl_->fixed_rlow;
    return impl_->fixed && impl_->fixed_rtop ? kRows : impl_->tuned_rlow;
}

void FusedFrame::setInterior(int mode) {
    if (mode < 0 || mode > 2)
        throw invalid_argument("fused frame: interior tile body must be 0 (autotune), 1 (on) or 2 (off), got " +
                               std::to_string(mode));
    impl_->fixed_inner = mode;
}

int FusedFrame::interior() const {
    if (impl_->fixed_inner) return impl_->fixed_inner;
    return impl_->tuned ? (impl_->tuned_inner ? 1 : 2) : 0;
}

void FusedFrame::setRows(int rows) {
    if (rows != 0 && rows != 4 && rows != 8)
        throw invalid_argument("fused frame: rows per thread must be 0 (autotune), 4 or 8, got " +
                               std::to_string(rows));
    impl_->fixed_rlow = rows;
}

unsigned long long FusedFrame::timeouts() const { return impl_->watch->read(); }

const std::shared_ptr<WaitWatch>& FusedFrame::watch() const { return impl_->watch; }

void FusedFrame::setWaitTest(int spins) {
    if (spins < 0) throw invalid_argument("fused frame: wait test spins must be >= 0");
    impl_->test_spins = spins;
}

int FusedFrame::waitTest() const { return impl_->test_spins; }

// ------------------------------------------------------------- WaitWatch

WaitWatch::~WaitWatch() {
    if (host) (void)hipHostFree(host);
}

unsigned long long WaitWatch::read() {
    if (!pending || !word) return total;
    if (!host) FF_HIP(hipHostMalloc((void**)&host, sizeof(unsigned), hipHostMallocDefault));
    FF_HIP(hipMemcpyAsync(host, word, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    FF_HIP(hipStreamSynchronize(stream));
    total = base + *host;
    pending = false;
    return total;
}

void WaitWatch::retire(const unsigned* next) {
    base = total;
    word = next;
}

void WaitCheck::follow(const std::shared_ptr<WaitWatch>& w) {
    if (watch == w) return;
    watch = w;
    seen = w ? w->total : 0;
}

void WaitCheck::check(const char* who) {
    if (!watch) return;
    const unsigned long long n = watch->read();
    if (n > seen) {
        const unsigned long long d = n - seen;
        seen = n;
        throw runtime_error(std::string(who) + ": " + std::to_string(d) +
                            " wait(s) of the fused frame kernel for the level above gave up after the spin limit "
                            "since the last check; the flow of those frames is not valid");
    }
}

}  // namespace ffamd
