// ff_oracle.hpp -- CPU restatement of the Adarve-Mahony optical-flow filter
// hot path (flowfilter::gpu), written from the reference kernel sources.
//
// TEST INFRASTRUCTURE ONLY. This header is the checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load the shared library built from it. The product path never calls it.
//
// Templated on the arithmetic type T:
//   T = float  : the parity oracle. Same operations, same order, same
//                rounding points as the reference kernels (policy below).
//   T = double : the FP64 oracle. Same algorithm evaluated in double;
//                the u8 pyramid and the upper-level flow index are taken
//                in exact integer arithmetic (their mathematically intended
//                values).
//
// Floating-point policy (applied identically in the HIP kernels):
//   * Built with -ffp-contract=off; no implicit contraction anywhere.
//   * Every accumulation of the form `acc += a*b` / `acc -= a*b` written in
//     one statement in the reference kernels is one fused multiply-add
//     (mad() below), which is what nvcc's default -fmad=true emits for them.
//     Products stored in named variables and reused (ax2, ay2, N01, dt*Ud)
//     are rounded on their own.
//   * Division is IEEE (correctly rounded); min/max are fminf/fmaxf
//     (a NaN operand is ignored).
//   * CUDA texture reads (clamp address mode, point filter, unnormalized
//     coordinates) are clamp-to-edge index reads. An unsigned-char texture
//     read with cudaReadModeNormalizedFloat is defined as (float)v / 255.0f.
//     The hardware conversion itself is not pinned by anything in the
//     reference; this definition is the one the GPU build implements too.
//
// Reference files (relative to the reference repository root):
//   src/gpu/device/imagemodel_k.cu, propagation_k.cu, update_k.cu,
//   flowsmoothing_k.cu, pyramid_k.cu, misc_k.cu  (per-pixel arithmetic)
//   src/gpu/{imagemodel,propagation,update,flowsmoothing,pyramid,
//   flowfilter}.cu                                (stage sequencing)
//   src/gpu/image.cu:228-356                      (texture modes)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ffo {

// ---------------------------------------------------------------- fields

// Dense row-major h x w x c array (no pitch; the oracle has no alignment
// concerns).
template <class T>
struct Field {
    int h = 0, w = 0, c = 1;
    std::vector<T> v;
    Field() = default;
    Field(int h_, int w_, int c_) : h(h_), w(w_), c(c_), v((size_t)h_ * w_ * c_, T(0)) {}
    T& at(int y, int x, int k = 0) { return v[((size_t)y * w + x) * c + k]; }
    T at(int y, int x, int k = 0) const { return v[((size_t)y * w + x) * c + k]; }
    // Clamp-to-edge point read: cudaAddressModeClamp, cudaFilterModePoint,
    // unnormalized coordinates (src/gpu/image.cu:228-231, 324-328).
    T tex(int y, int x, int k = 0) const {
        y = y < 0 ? 0 : (y >= h ? h - 1 : y);
        x = x < 0 ? 0 : (x >= w ? w - 1 : x);
        return at(y, x, k);
    }
    void zero() { std::fill(v.begin(), v.end(), T(0)); }
};

// Planted mistakes, for the tests that show each pinning check fails on a
// wrong oracle (0, the default, is the faithful restatement). Each one is a
// plausible transcription error of the reference:
//   MUT_SMOOTHED_TOP_FEED  level L-2 reads the top level's smoothed flow
//                          instead of its unsmoothed one (flowfilter.cu:663-672)
//   MUT_DESCENDING_UPDATE  lower levels update from L-2 down to 0 instead of
//                          0 up to L-2 (flowfilter.cu:710-712)
//   MUT_TOP_UPDATES_LAST   the top level updates after the lower levels
//                          (flowfilter.cu:708)
//   MUT_PYR_TAP_SHIFT      pyramid taps at 2i, 2i+1, 2i+2 instead of 2i-1,
//                          2i, 2i+1 (pyramid_k.cu:33-35, 67-69, 100-102,
//                          134-136)
//   MUT_PYR_SUM_DOUBLE     the float pyramid adds its side taps in double
//                          instead of float (pyramid_k.cu:105, 139)
enum Mutation {
    MUT_NONE = 0,
    MUT_SMOOTHED_TOP_FEED = 1,
    MUT_DESCENDING_UPDATE = 2,
    MUT_TOP_UPDATES_LAST = 3,
    MUT_PYR_TAP_SHIFT = 4,
    MUT_PYR_SUM_DOUBLE = 5,
};
inline int& mutation() { static int m = MUT_NONE; return m; }

inline float mad(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double mad(double a, double b, double c) { return std::fma(a, b, c); }
inline float fmin_(float a, float b) { return std::fminf(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmaxf(a, b); }
inline double fmax_(double a, double b) { return std::fmax(a, b); }

// Normalized read of an 8-bit texel.
template <class T> inline T unorm8(uint8_t v) { return T(v) / T(255); }

template <class T> inline T read_px(const Field<uint8_t>& img, int y, int x) {
    return unorm8<T>(img.tex(y, x));
}
template <class T> inline T read_px(const Field<float>& img, int y, int x) {
    return T(img.tex(y, x));
}

// Number of OpenMP threads used by the row loops (1 = serial). Results are
// identical for any thread count: every output pixel is computed
// independently.
inline int& num_threads() { static int n = 1; return n; }

#define FFO_ROWS_PARALLEL _Pragma("omp parallel for schedule(static) num_threads(ffo::num_threads())")

// ----------------------------------------------------------- image model

// Masks, imagemodel_k.cu:21-22.
template <class T> inline T smooth_mask(int i) {
    static const T m[5] = {T(0.0625), T(0.25), T(0.375), T(0.25), T(0.0625)};
    return m[i];
}
template <class T> inline T diff_mask(int i) {
    static const T m[5] = {T(-0.125), T(-0.25), T(0), T(0.25), T(0.125)};
    return m[i];
}

// imagePrefilter_k (imagemodel_k.cu:31-70) followed by imageModel_k
// (imagemodel_k.cu:76-133). The prefiltered pair is stored as
// {smooth_y, smooth_x} (:69): component 0 is the Y-smoothed value,
// component 1 the X-smoothed value.
template <class T, class Tin>
void image_model(const Field<Tin>& img, Field<T>& a0, Field<T>& grad) {
    const int h = img.h, w = img.w;
    Field<T> pre(h, w, 2);
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T sx = T(0), sy = T(0);
            for (int c = -2; c <= 2; ++c) sx = mad(smooth_mask<T>(c + 2), read_px<T>(img, y, x + c), sx);
            for (int r = -2; r <= 2; ++r) sy = mad(smooth_mask<T>(r + 2), read_px<T>(img, y + r, x), sy);
            pre.at(y, x, 0) = sy;
            pre.at(y, x, 1) = sx;
        }
    a0 = Field<T>(h, w, 1);
    grad = Field<T>(h, w, 2);
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T dx = T(0), dy = T(0), sm = T(0);
            for (int c = -2; c <= 2; ++c) {
                const T e = pre.tex(y, x + c, 0);
                dx = mad(diff_mask<T>(c + 2), e, dx);
                sm = mad(smooth_mask<T>(c + 2), e, sm);
            }
            for (int r = -2; r <= 2; ++r) dy = mad(diff_mask<T>(r + 2), pre.tex(y + r, x, 1), dy);
            grad.at(y, x, 0) = dx;
            grad.at(y, x, 1) = dy;
            a0.at(y, x) = sm;
        }
}

// ----------------------------------------------------------- propagation

// One upwind sub-pass along x (axis 0) or y (axis 1) of the flow and of
// optional payloads advected with the same dominant velocity.
//   flowPropagateX_k    propagation_k.cu:16-72
//   flowPropagateY_k    propagation_k.cu:75-131
//   flowPropagatePayloadX_k / Y_k  propagation_k.cu:134-334
// For each component k:  out = z - (dt*Ud) * (Ud >= 0 ? z - m : p - z),
// with Ud = (|p.u| - |m.u| > 0) ? p.u : m.u, u the velocity component of
// the pass. Pixels closer than `border` to any image edge keep z.
template <class T>
void propagate_pass(int axis, const Field<T>& flow, Field<T>& flow_out, const Field<T>* scalar,
                    Field<T>* scalar_out, const Field<T>* vec, Field<T>* vec_out, T dt, int border) {
    const int h = flow.h, w = flow.w;
    flow_out = Field<T>(h, w, 2);
    if (scalar) *scalar_out = Field<T>(h, w, 1);
    if (vec) *vec_out = Field<T>(h, w, 2);
    const int dy = axis == 1 ? 1 : 0, dx = axis == 0 ? 1 : 0;
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const T mu = flow.tex(y - dy, x - dx, axis), pu = flow.tex(y + dy, x + dx, axis);
            const T central = std::fabs(pu) - std::fabs(mu);
            const T Ud = central > T(0) ? pu : mu;
            const T dtUd = dt * Ud;
            const bool back = Ud >= T(0);
            const bool in = (x >= border && x < w - border) && (y >= border && y < h - border);
            auto step = [&](const Field<T>& f, int k) -> T {
                const T m = f.tex(y - dy, x - dx, k), z = f.at(y, x, k), p = f.tex(y + dy, x + dx, k);
                const T dm = z - m, dp = p - z;
                const T r = mad(-dtUd, back ? dm : dp, z);
                return in ? r : z;
            };
            flow_out.at(y, x, 0) = step(flow, 0);
            flow_out.at(y, x, 1) = step(flow, 1);
            if (scalar) scalar_out->at(y, x) = step(*scalar, 0);
            if (vec) {
                vec_out->at(y, x, 0) = step(*vec, 0);
                vec_out->at(y, x, 1) = step(*vec, 1);
            }
        }
}

// FlowPropagator::compute (propagation.cu:80-134): X then Y, `iterations`
// times; dt = 1/iterations (:147). With invert, the input is first scaled
// by -1.0f (scalarProductF2_k, misc_k.cu:18-39).
template <class T>
Field<T> propagate(const Field<T>& flow_in, int iterations, int border, bool invert) {
    const T dt = T(1) / T(iterations);
    Field<T> src = flow_in;
    if (invert)
        for (auto& v : src.v) v = T(-1) * v;
    Field<T> bx, by;
    propagate_pass<T>(0, src, bx, nullptr, nullptr, nullptr, nullptr, dt, border);
    propagate_pass<T>(1, bx, by, nullptr, nullptr, nullptr, nullptr, dt, border);
    for (int n = 0; n < iterations - 1; ++n) {
        propagate_pass<T>(0, by, bx, nullptr, nullptr, nullptr, nullptr, dt, border);
        propagate_pass<T>(1, bx, by, nullptr, nullptr, nullptr, nullptr, dt, border);
    }
    return by;
}

// FlowPropagatorPayload::compute (propagation.cu:324-373).
template <class T>
void propagate_payload(const Field<T>& flow_in, const Field<T>& scalar_in, const Field<T>& vec_in,
                       int iterations, int border, Field<T>& flow_out, Field<T>& scalar_out,
                       Field<T>& vec_out) {
    const T dt = T(1) / T(iterations);
    Field<T> fx, sx, vx;
    propagate_pass<T>(0, flow_in, fx, &scalar_in, &sx, &vec_in, &vx, dt, border);
    propagate_pass<T>(1, fx, flow_out, &sx, &scalar_out, &vx, &vec_out, dt, border);
    for (int n = 0; n < iterations - 1; ++n) {
        propagate_pass<T>(0, flow_out, fx, &scalar_out, &sx, &vec_out, &vx, dt, border);
        propagate_pass<T>(1, fx, flow_out, &sx, &scalar_out, &vx, &vec_out, dt, border);
    }
}

// ---------------------------------------------------------------- update

template <class T> inline bool bad_(T v) { return std::isinf(v) || std::isnan(v); }

// Closed-form regularised 2x2 solve shared by flowUpdate_k and
// deltaFlowUpdate_k (update_k.cu:43-63, 114-134):
//   (gamma I + g g^T)^-1 (gamma f + g Yt)  via the adjugate.
template <class T>
inline void solve2x2(T gx, T gy, T Yt, T fx, T fy, T gamma, T& ox, T& oy) {
    const T ax2 = gx * gx;
    const T ay2 = gy * gy;
    const T N00 = gamma + ay2;
    const T N01 = -gx * gy;
    const T N10 = N01;
    const T N11 = gamma + ax2;
    const T rdet = T(1) / (gamma * ((gamma + ax2) + ay2));
    const T qx = mad(gamma, fx, gx * Yt);
    const T qy = mad(gamma, fy, gy * Yt);
    ox = mad(N00, qx, N01 * qy) * rdet;
    oy = mad(N10, qx, N11 * qy) * rdet;
}

// flowUpdate_k (update_k.cu:15-79): solve, clamp to +-maxflow with
// fminf/fmaxf, then NaN/Inf -> 0 (which the clamp already makes
// unreachable for finite maxflow: NaN -> +maxflow, Inf -> +-maxflow).
template <class T>
void update(const Field<T>& a0, const Field<T>& grad, const Field<T>& a0old, const Field<T>& flow_old,
            T gamma, T maxflow, Field<T>& flow_out, Field<T>& image_out) {
    const int h = a0.h, w = a0.w;
    Field<T> fo(h, w, 2), io(h, w, 1);
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const T gx = grad.at(y, x, 0), gy = grad.at(y, x, 1);
            const T a = a0.at(y, x);
            const T Yt = a0old.at(y, x) - a;
            T ox, oy;
            solve2x2<T>(gx, gy, Yt, flow_old.at(y, x, 0), flow_old.at(y, x, 1), gamma, ox, oy);
            ox = fmax_(-maxflow, fmin_(ox, maxflow));
            oy = fmax_(-maxflow, fmin_(oy, maxflow));
            ox = bad_(ox) ? T(0) : ox;
            oy = bad_(oy) ? T(0) : oy;
            fo.at(y, x, 0) = ox;
            fo.at(y, x, 1) = oy;
            io.at(y, x) = a;
        }
    flow_out = std::move(fo);
    image_out = std::move(io);
}

// Index of the upper-level texel read by deltaFlowUpdate_k through a
// normalized-coordinate, point-filtered, clamped texture
// (update_k.cu:151-155, update.cu:308-313):
//   u = (float)x / (float)(w-1);  i = min(floor(u * Wu), Wu-1).
// FP32: evaluated in float exactly as written. FP64: the exact value
// floor(x * Wu / (w-1)) in integers. A size of 1 maps to index 0.
inline int upper_index(int x, int w, int wu, float) {
    if (w <= 1) return 0;
    const float u = (float)x / (float)(w - 1);
    const float s = u * (float)wu;
    int i = (int)std::floor(s);
    if (i > wu - 1) i = wu - 1;
    if (i < 0) i = 0;
    return i;
}
inline int upper_index(int x, int w, int wu, double) {
    if (w <= 1) return 0;
    long long i = ((long long)x * wu) / (w - 1);
    if (i > wu - 1) i = wu - 1;
    return (int)i;
}

// deltaFlowUpdate_k (update_k.cu:82-173): solve on the delta flow,
// sanitize (NaN/Inf -> 0), clamp +-0.5*maxflow, add 2x the upper-level flow
// (point sampled), clamp +-maxflow.
template <class T>
void delta_update(const Field<T>& a0, const Field<T>& grad, const Field<T>& a0old, const Field<T>& dflow_old,
                  const Field<T>& upper, T gamma, T maxflow, Field<T>& dflow_out, Field<T>& flow_out,
                  Field<T>& image_out) {
    const int h = a0.h, w = a0.w;
    Field<T> d(h, w, 2), f(h, w, 2), io(h, w, 1);
    const T half = T(0.5) * maxflow;
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y) {
        const int iy = upper_index(y, h, upper.h, T(0));
        for (int x = 0; x < w; ++x) {
            const int ix = upper_index(x, w, upper.w, T(0));
            const T gx = grad.at(y, x, 0), gy = grad.at(y, x, 1);
            const T a = a0.at(y, x);
            const T Yt = a0old.at(y, x) - a;
            T ox, oy;
            solve2x2<T>(gx, gy, Yt, dflow_old.at(y, x, 0), dflow_old.at(y, x, 1), gamma, ox, oy);
            ox = bad_(ox) ? T(0) : ox;
            oy = bad_(oy) ? T(0) : oy;
            ox = fmax_(-half, fmin_(ox, half));
            oy = fmax_(-half, fmin_(oy, half));
            const T ux = T(2) * upper.at(iy, ix, 0), uy = T(2) * upper.at(iy, ix, 1);
            T fx = ox + ux, fy = oy + uy;
            fx = fmax_(-maxflow, fmin_(fx, maxflow));
            fy = fmax_(-maxflow, fmin_(fy, maxflow));
            d.at(y, x, 0) = ox;
            d.at(y, x, 1) = oy;
            f.at(y, x, 0) = fx;
            f.at(y, x, 1) = fy;
            io.at(y, x) = a;
        }
    }
    dflow_out = std::move(d);
    flow_out = std::move(f);
    image_out = std::move(io);
}

// ------------------------------------------------------------- smoothing

// flowSmoothX_k / flowSmoothY_k (flowsmoothing_k.cu:21-87): 5-tap box with
// weight 0.2, accumulated from 0 for offsets -2..2, clamp-to-edge.
template <class T>
void smooth_pass(int axis, const Field<T>& in, Field<T>& out) {
    const int h = in.h, w = in.w;
    out = Field<T>(h, w, 2);
    const int dy = axis == 1 ? 1 : 0, dx = axis == 0 ? 1 : 0;
    const T c5 = T(0.2f);  // __constant__ float flowSmooth5_k[] = {0.2, ...}
    FFO_ROWS_PARALLEL
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T s0 = T(0), s1 = T(0);
            for (int c = -2; c <= 2; ++c) {
                s0 = mad(c5, in.tex(y + c * dy, x + c * dx, 0), s0);
                s1 = mad(c5, in.tex(y + c * dy, x + c * dx, 1), s1);
            }
            out.at(y, x, 0) = s0;
            out.at(y, x, 1) = s1;
        }
}

// FlowSmoother::compute (flowsmoothing.cu:74-104).
template <class T>
Field<T> smooth(const Field<T>& in, int iterations) {
    Field<T> bx, by;
    smooth_pass<T>(0, in, bx);
    smooth_pass<T>(1, bx, by);
    for (int n = 0; n < iterations - 1; ++n) {
        smooth_pass<T>(0, by, bx);
        smooth_pass<T>(1, bx, by);
    }
    return by;
}

// --------------------------------------------------------------- pyramid

// imageDownX_uint8_k / imageDownY_uint8_k (pyramid_k.cu:14-78):
//   v = (float)(0.5*(double)t0 + 0.25*(double)(tm + tp)),  tm + tp in float,
//   out = (unsigned char)(255 * v)   (float product, truncation).
// The taps are the normalized reads at 2i-1, 2i, 2i+1 (clamped).
// FP64 oracle: the exact value floor((2*k0 + km + kp) / 4).
inline uint8_t down_u8(uint8_t km, uint8_t k0, uint8_t kp, float) {
    const float tm = unorm8<float>(km), t0 = unorm8<float>(k0), tp = unorm8<float>(kp);
    const float s = tm + tp;
    const float v = (float)(0.5 * (double)t0 + 0.25 * (double)s);
    return (uint8_t)(255 * v);
}
inline uint8_t down_u8(uint8_t km, uint8_t k0, uint8_t kp, double) {
    return (uint8_t)((2 * (int)k0 + (int)km + (int)kp) / 4);
}

// imageDownX_float_k / imageDownY_float_k (pyramid_k.cu:81-145): the side
// taps are added in float, the rest in double (0.5 and 0.25 are double
// literals), then rounded to float.
inline float down_f32(float tm, float t0, float tp, float) {
    if (mutation() == MUT_PYR_SUM_DOUBLE) return (float)(0.5 * (double)t0 + 0.25 * ((double)tm + (double)tp));
    const float s = tm + tp;
    return (float)(0.5 * (double)t0 + 0.25 * (double)s);
}
inline float down_f32(float tm, float t0, float tp, double) {
    return (float)(0.5 * (double)t0 + 0.25 * ((double)tm + (double)tp));
}

// ImagePyramid::compute (pyramid.cu:110-167); level sizes by integer
// halving, X then Y (pyramid.cu:74,83). Level 0 is the input itself.
template <class T, class Tin>
std::vector<Field<Tin>> pyramid(const Field<Tin>& img, int levels) {
    std::vector<Field<Tin>> out;
    out.push_back(img);
    // first of the three taps: 2i - 1 (MUT_PYR_TAP_SHIFT: 2i)
    const int t0 = mutation() == MUT_PYR_TAP_SHIFT ? 0 : -1;
    for (int l = 1; l < levels; ++l) {
        const Field<Tin>& src = out.back();
        const int h = src.h, w2 = src.w / 2;
        Field<Tin> fx(h, w2, 1);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w2; ++x) {
                const int c = 2 * x + t0;
                if constexpr (sizeof(Tin) == 1)
                    fx.at(y, x) = down_u8(src.tex(y, c), src.tex(y, c + 1), src.tex(y, c + 2), T(0));
                else
                    fx.at(y, x) = down_f32(src.tex(y, c), src.tex(y, c + 1), src.tex(y, c + 2), T(0));
            }
        const int h2 = h / 2;
        Field<Tin> fy(h2, w2, 1);
        for (int y = 0; y < h2; ++y)
            for (int x = 0; x < w2; ++x) {
                const int r = 2 * y + t0;
                if constexpr (sizeof(Tin) == 1)
                    fy.at(y, x) = down_u8(fx.tex(r, x), fx.tex(r + 1, x), fx.tex(r + 2, x), T(0));
                else
                    fy.at(y, x) = down_f32(fx.tex(r, x), fx.tex(r + 1, x), fx.tex(r + 2, x), T(0));
            }
        out.push_back(std::move(fy));
    }
    return out;
}

// --------------------------------------------------------------- display

// Texel read by a normalised-coordinate, point-filtered, clamped texture
// (the colour wheel of FlowToColor::configure, display.cu:59-67) at
// coordinate c on an axis of n texels: floor(c * n) clamped to [0, n-1]; a
// NaN coordinate reads texel 0 (the texture unit's own behaviour for it is
// not pinned by the reference).
inline int wheel_texel_index(float c, int n) {
    const float s = c * (float)n;
    if (!(s > 0.0f)) return 0;  // negative, zero or NaN
    if (s >= (float)(n - 1)) return n - 1;
    return (int)std::floor(s);
}

// flowToColor_k (display_k.cu:14-43): out (h x w x 4) = wheel texel at
// ((u + m) / (2m), (v + m) / (2m)), the normalisation in float with IEEE
// division; wheel is wh x ww x 4 u8, row-major.
inline void flow_to_color(const Field<float>& flow, float maxflow, const Field<uint8_t>& wheel,
                          Field<uint8_t>& out) {
    out = Field<uint8_t>(flow.h, flow.w, 4);
    const float den = 2.0f * maxflow;
    for (int y = 0; y < flow.h; ++y)
        for (int x = 0; x < flow.w; ++x) {
            const float cu = (flow.at(y, x, 0) + maxflow) / den;
            const float cv = (flow.at(y, x, 1) + maxflow) / den;
            const int ix = wheel_texel_index(cu, wheel.w), iy = wheel_texel_index(cv, wheel.h);
            for (int k = 0; k < 4; ++k) out.at(y, x, k) = wheel.at(iy, ix, k);
        }
}

// ------------------------------------- rotational flow and Lax-Wendroff

// perspectiveCamera and createPerspectiveCamera (camera.h:20-25,
// camera.cu:14-29).
struct Camera {
    float alphaX, alphaY, centerX, centerY;
};
inline Camera create_perspective_camera(float focal, int h, int w, float sensor_h, float sensor_w) {
    const int ix = w - 1, iy = h - 1;
    return Camera{focal * (float)ix / sensor_w, focal * (float)iy / sensor_h, 0.5f * (float)ix, 0.5f * (float)iy};
}

// rotationalOpticalFlow_k (rotation_k.cu:17-43): p = ((x - cx)/ax,
// (y - cy)/ay, 1) (camera_k.h:22-27), c = w x p (math_k.h:21-26), flow =
// (ax c.x + (cx - x) c.z, ay c.y + (cy - y) c.z). Contraction policy (the
// GPU kernel's too): in a*b - c*d and a*b + c*d the first product is fused.
inline void rotational_flow(const Camera& cam, float wx, float wy, float wz, int h, int w, Field<float>& flow) {
    flow = Field<float>(h, w, 2);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float fx = (float)x, fy = (float)y;
            const float px = (fx - cam.centerX) / cam.alphaX, py = (fy - cam.centerY) / cam.alphaY, pz = 1.0f;
            const float c0 = std::fmaf(wy, pz, -(wz * py));
            const float c1 = std::fmaf(wz, px, -(wx * pz));
            const float c2 = std::fmaf(wx, py, -(wy * px));
            flow.at(y, x, 0) = std::fmaf(cam.alphaX, c0, (cam.centerX - fx) * c2);
            flow.at(y, x, 1) = std::fmaf(cam.alphaY, c1, (cam.centerY - fy) * c2);
        }
}

// One Lax-Wendroff pass (LaxWendroffX_k axis 0 / LaxWendroffY_k axis 1,
// propagation_k.h:49-132) of an image of c float channels (1 or 4):
// out = z - R/2 (p - m) + R^2/2 (m - 2z + p), R = dt * flow[axis], neighbours
// clamped. One channel: the update in double (the reference's 0.5 is a
// double literal), the second difference in float along x (2 * z) and in
// double along y (2.0 * z); four channels: everything in float (the
// reference's float4 operators), with R/2 and R^2/2 rounded to float. The
// products that feed the two updates are fused.
inline void lax_wendroff_pass(int axis, const Field<float>& flow, const Field<float>& in, Field<float>& out, float dt) {
    const int h = in.h, w = in.w, c = in.c;
    out = Field<float>(h, w, c);
    const int dx = axis == 0 ? 1 : 0, dy = axis == 1 ? 1 : 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float R = dt * flow.at(y, x, axis);
            for (int k = 0; k < c; ++k) {
                const float m = in.tex(y - dy, x - dx, k), z = in.at(y, x, k), p = in.tex(y + dy, x + dx, k);
                const float d0 = p - m;
                if (c == 1) {
                    const double d2 = axis == 0 ? (double)((m - 2.0f * z) + p)
                                                : ((double)m - 2.0 * (double)z) + (double)p;
                    const double hR = 0.5 * (double)R;
                    const double t = std::fma(-hR, (double)d0, (double)z);
                    out.at(y, x, k) = (float)std::fma(hR * (double)R, d2, t);
                } else {
                    const float d2 = (m - 2.0f * z) + p;
                    const float hR = (float)(0.5 * (double)R), kR = (float)(0.5 * (double)R * (double)R);
                    out.at(y, x, k) = std::fmaf(kR, d2, std::fmaf(-hR, d0, z));
                }
            }
        }
}

// LaxWendroffPropagator::compute (propagation.cu:572-644): a Y pass then an X
// pass, max(iterations, 1) times; dt = 1/iterations, 0 when iterations is 0
// (never set).
inline Field<float> lax_wendroff(const Field<float>& flow, const Field<float>& img, int iterations) {
    const float dt = iterations > 0 ? 1.0f / float(iterations) : 0.0f;
    Field<float> by, bx;
    lax_wendroff_pass(1, flow, img, by, dt);
    lax_wendroff_pass(0, flow, by, bx, dt);
    for (int n = 0; n < iterations - 1; ++n) {
        lax_wendroff_pass(1, flow, bx, by, dt);
        lax_wendroff_pass(0, flow, by, bx, dt);
    }
    return bx;
}

// --------------------------------------------------------------- filters

// Single-level filter state and recurrence, FlowFilter (flowfilter.cu:23-330).
// Also used as the top level of the pyramid. Input element type Tin is
// uint8_t or float (FlowFilter::setInputImage accepts item size 1 or 4,
// flowfilter.cu:200-216).
template <class T>
struct TopLevel {
    int h = 0, w = 0;
    bool u8 = true;
    T gamma = T(1), maxflow = T(1);  // FlowUpdate defaults (update.h:29-30)
    int prop_iters = 1, smooth_iters = 1, border = 3;
    bool first = true;
    Field<T> a0, grad, prop_flow, flow_upd, image_upd, smoothed;
    void init(int h_, int w_, bool u8_) {
        h = h_; w = w_; u8 = u8_;
        prop_flow = Field<T>(h, w, 2);
        flow_upd = Field<T>(h, w, 2);
        image_upd = Field<T>(h, w, 1);
        smoothed = Field<T>(h, w, 2);
        a0 = Field<T>(h, w, 1);
        grad = Field<T>(h, w, 2);
        first = true;
    }
    // FlowFilter::setGamma (flowfilter.cu:276-285): u8 input scales by 1/255^2.
    void set_gamma(float g) {
        if (u8) gamma = T(g / (255.0f * 255.0f)); else gamma = T(g);
    }
    // FlowFilter::setMaxFlow (flowfilter.cu:293-296).
    void set_maxflow(float m) { maxflow = T(m); prop_iters = int(std::ceil(m)); }
    template <class Tin> void compute_image_model(const Field<Tin>& img) { image_model<T>(img, a0, grad); }
    void compute_propagation() { prop_flow = propagate<T>(smoothed, prop_iters, border, false); }
    // FlowFilter::computeUpdate (flowfilter.cu:175-197).
    void compute_update() {
        if (first) { image_upd = a0; first = false; }
        Field<T> img_new;
        update<T>(a0, grad, image_upd, prop_flow, gamma, maxflow, flow_upd, img_new);
        image_upd = std::move(img_new);
        smoothed = smooth<T>(flow_upd, smooth_iters);
    }
    // FlowFilter::compute (flowfilter.cu:125-153).
    template <class Tin> void compute(const Field<Tin>& img) {
        compute_image_model(img);
        if (first) { image_upd = a0; first = false; }
        compute_propagation();
        Field<T> img_new;
        update<T>(a0, grad, image_upd, prop_flow, gamma, maxflow, flow_upd, img_new);
        image_upd = std::move(img_new);
        smoothed = smooth<T>(flow_upd, smooth_iters);
    }
};

// Non-top pyramid level, DeltaFlowFilter (flowfilter.cu:338-615).
template <class T>
struct DeltaLevel {
    int h = 0, w = 0;
    bool u8 = true;
    T gamma = T(1), maxflow = T(1);
    int prop_iters = 1, smooth_iters = 1, border = 3;
    bool first = true;
    Field<T> a0, grad, prop_flow, prop_scalar, prop_vec, dflow_upd, flow_upd, image_upd, smoothed;
    void init(int h_, int w_, bool u8_) {
        h = h_; w = w_; u8 = u8_;
        a0 = Field<T>(h, w, 1);
        grad = Field<T>(h, w, 2);
        prop_flow = Field<T>(h, w, 2);
        prop_scalar = Field<T>(h, w, 1);
        prop_vec = Field<T>(h, w, 2);
        dflow_upd = Field<T>(h, w, 2);
        flow_upd = Field<T>(h, w, 2);
        image_upd = Field<T>(h, w, 1);
        smoothed = Field<T>(h, w, 2);
        first = true;
    }
    void set_gamma(float g) {
        if (u8) gamma = T(g / (255.0f * 255.0f)); else gamma = T(g);
    }
    void set_maxflow(float m) { maxflow = T(m); prop_iters = int(std::ceil(m)); }
    template <class Tin> void compute_image_model(const Field<Tin>& img) { image_model<T>(img, a0, grad); }
    // DeltaFlowFilter::computePropagation: FlowPropagatorPayload with the
    // smoothed flow as velocity, the updated image and delta flow as
    // payloads (flowfilter.cu:401-407).
    void compute_propagation() {
        propagate_payload<T>(smoothed, image_upd, dflow_upd, prop_iters, border, prop_flow, prop_scalar,
                             prop_vec);
    }
    // DeltaFlowFilter::computeUpdate (flowfilter.cu:480-503): on the first
    // frame both the old image and the propagated scalar payload take a0.
    void compute_update(const Field<T>& upper) {
        if (first) { image_upd = a0; prop_scalar = a0; first = false; }
        run_update(upper);
    }
    // DeltaFlowFilter::compute (flowfilter.cu:429-457), the standalone
    // filter: on the first frame only the old image takes a0, before the
    // propagation, which then carries it as the scalar payload.
    template <class Tin> void compute(const Field<Tin>& img, const Field<T>& upper) {
        compute_image_model(img);
        if (first) { image_upd = a0; first = false; }
        compute_propagation();
        run_update(upper);
    }
    void run_update(const Field<T>& upper) {
        delta_update<T>(a0, grad, prop_scalar, prop_vec, upper, gamma, maxflow, dflow_upd, flow_upd, image_upd);
        smoothed = smooth<T>(flow_upd, smooth_iters);
    }
};

// PyramidalFlowFilter (flowfilter.cu:621-897). Input is always uint8
// (flowfilter.cu:652).
template <class T>
struct Pyramidal {
    int h = 0, w = 0, L = 1;
    TopLevel<T> top;
    std::vector<DeltaLevel<T>> low;  // low[l] = level l, l = 0 .. L-2
    void init(int h_, int w_, int levels) {
        if (levels <= 0) throw std::invalid_argument("levels should be greater than zero");
        h = h_; w = w_; L = levels;
        int hh = h, ww = w;
        std::vector<std::pair<int, int>> sz;
        sz.push_back({hh, ww});
        for (int l = 1; l < L; ++l) { ww /= 2; hh /= 2; sz.push_back({hh, ww}); }
        top.init(sz[L - 1].first, sz[L - 1].second, true);
        low.resize(L - 1);
        for (int l = 0; l < L - 1; ++l) low[l].init(sz[l].first, sz[l].second, true);
    }
    void set_gamma(int level, float g) { if (level == L - 1) top.set_gamma(g); else low[level].set_gamma(g); }
    void set_smooth_iterations(int level, int n) {
        if (level == L - 1) top.smooth_iters = n; else low[level].smooth_iters = n;
    }
    // PyramidalFlowFilter::setMaxFlow (flowfilter.cu:849-866).
    void set_maxflow(float m) {
        if (L == 1) { top.set_maxflow(m); return; }
        float ml = m;
        for (int l = 0; l < L - 1; ++l) { low[l].set_maxflow(ml); ml /= 2.0f; }
        top.set_maxflow(ml);
    }
    void set_border(int b) { top.border = b; for (auto& d : low) d.border = b; }
    // Upper-level flow read by level l: the top level's *unsmoothed*
    // updated flow for l = L-2, else level l+1's smoothed flow
    // (flowfilter.cu:663-672).
    const Field<T>& upper_of(int l) const {
        if (l == L - 2) return mutation() == MUT_SMOOTHED_TOP_FEED ? top.smoothed : top.flow_upd;
        return low[l + 1].smoothed;
    }
    // PyramidalFlowFilter::compute (flowfilter.cu:685-716): image models and
    // propagation of every level, then the updates: the top level first
    // (:708), then levels 0, 1, ..., L-2 (:710-712), so that level l < L-2
    // reads level l+1's smoothed flow of the previous frame.
    void compute(const Field<uint8_t>& img) {
        std::vector<Field<uint8_t>> pyr = pyramid<T, uint8_t>(img, L);
        if (L == 1) { top.compute(pyr[0]); return; }
        top.compute_image_model(pyr[L - 1]);
        top.compute_propagation();
        for (int l = 0; l < L - 1; ++l) { low[l].compute_image_model(pyr[l]); low[l].compute_propagation(); }
        const int m = mutation();
        if (m != MUT_TOP_UPDATES_LAST) top.compute_update();
        if (m == MUT_DESCENDING_UPDATE)
            for (int l = L - 2; l >= 0; --l) low[l].compute_update(upper_of(l));
        else
            for (int l = 0; l < L - 1; ++l) low[l].compute_update(upper_of(l));
        if (m == MUT_TOP_UPDATES_LAST) top.compute_update();
    }
    // downloadFlow (flowfilter.cu:734-741): smoothed flow of level 0.
    const Field<T>& flow() const { return L == 1 ? top.smoothed : low[0].smoothed; }
    // downloadImage (flowfilter.cu:744-751).
    const Field<T>& image() const { return L == 1 ? top.image_upd : low[0].image_upd; }
    // getFlow (flowfilter.cu:718-725): the top level's unsmoothed flow when
    // L == 1, level 0's smoothed flow otherwise.
    const Field<T>& device_flow() const { return L == 1 ? top.flow_upd : low[0].smoothed; }
};

}  // namespace ffo
