// ff_adjoint.hip -- differentiable form of the filter (gfx950): batched
// forward stage kernels on float images and their hand-written adjoints.
//
// Data: B images of H x W, contiguous (scalar fields [B][H][W] float,
// vector fields [B][H][W] float2). Forward kernels use the same per-pixel
// arithmetic as the stage kernels (ff_device.hpp), so a float-image filter
// built from them gives the same flow as the filter itself. Each backward
// kernel computes the transpose of its forward stage's Jacobian at the
// forward inputs (vector-Jacobian product); piecewise choices (upwind
// direction, clamps, the sanitising of non-finite values) are held at the
// forward values, as any derivative of the filter must.
//
// The backward kernels gather (each output of the adjoint sums the
// contributions that reach it), so they are deterministic and need no
// atomics. `acc` is a bit mask over a kernel's outputs (bit k: output k, in
// argument order, is added to the destination instead of written), so the
// sums of a chain rule land in place without extra passes over HBM. A null
// output-gradient input reads as zero.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ff_device.hpp"
#include "ff_kernels.hpp"
#include "ff_runtime.hpp"

namespace ffamd {

namespace {

constexpr int BX = 64, BY = 4;

struct Dims {
    int H, W;
};

__device__ __forceinline__ size_t at(int b, int y, int x, const Dims& d) {
    return ((size_t)b * d.H + y) * d.W + x;
}

inline dim3 grid3(int B, int H, int W) { return dim3((W + BX - 1) / BX, (H + BY - 1) / BY, B); }

// threadIdx.y in a workgroup 64 threads wide (every kernel here): one value
// per wave, so it is read into a scalar register, and the row indices,
// clamps and row conditions derived from it stay on the scalar unit (the
// compiler cannot tell that threadIdx.y is uniform across a wave).
__device__ __forceinline__ int wave_y() { return __builtin_amdgcn_readfirstlane((int)threadIdx.y); }

__device__ __forceinline__ void put(float* p, size_t i, float v, int acc) { p[i] = acc ? p[i] + v : v; }
__device__ __forceinline__ void put2(float2* p, size_t i, float2 v, int acc) {
    if (acc) {
        const float2 o = p[i];
        v = make_float2(o.x + v.x, o.y + v.y);
    }
    p[i] = v;
}

// Adjoint of a 5-tap correlation with clamped reads along one axis:
// y[i] = sum_c w[c] v[clamp(i + c)], c = -2..2, i in [0, n). Returns
// dv[j] = sum over (i, c) with clamp(i + c) == j of w[c] dy[i], where dy(i)
// reads the output adjoint at position i.
template <class Wf, class Dy>
__device__ __forceinline__ float corr5_adjoint(int j, int n, Wf w, Dy dy) {
    float acc = 0.0f;
    for (int i = j - 2; i <= j + 2; ++i) {
        if (i < 0 || i >= n) continue;
#pragma unroll
        for (int c = -2; c <= 2; ++c)
            if (clampi(i + c, 0, n - 1) == j) acc = fmaf(w(c + 2), dy(i), acc);
    }
    // (an output i reaches j only if |i - j| <= 2, clamped or not)
    return acc;
}

struct WS {
    __device__ __forceinline__ float operator()(int i) const { return smooth_w(i); }
};
struct WD {
    __device__ __forceinline__ float operator()(int i) const { return diff_w(i); }
};
struct WB {
    __device__ __forceinline__ float operator()(int) const { return kBox; }
};

// ------------------------------------------------------------ image model

// imagePrefilter_k + imageModel_k on a float image (imagemodel_k.cu:31-133).
__global__ void __launch_bounds__(BX* BY) k_ad_image_model(const float* I, float* a0, float2* g, Dims d) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    float dx = 0.0f, dy = 0.0f, sm = 0.0f;
#pragma unroll
    for (int c = -2; c <= 2; ++c) {
        const int xc = clampi(x + c, 0, d.W - 1);
        float sy = 0.0f;
#pragma unroll
        for (int r = -2; r <= 2; ++r) sy = fmaf(smooth_w(r + 2), I[at(b, clampi(y + r, 0, d.H - 1), xc, d)], sy);
        dx = fmaf(diff_w(c + 2), sy, dx);
        sm = fmaf(smooth_w(c + 2), sy, sm);
    }
#pragma unroll
    for (int r = -2; r <= 2; ++r) {
        const int yr = clampi(y + r, 0, d.H - 1);
        float sx = 0.0f;
#pragma unroll
        for (int c = -2; c <= 2; ++c) sx = fmaf(smooth_w(c + 2), I[at(b, yr, clampi(x + c, 0, d.W - 1), d)], sx);
        dy = fmaf(diff_w(r + 2), sx, dy);
    }
    a0[at(b, y, x, d)] = sm;
    g[at(b, y, x, d)] = make_float2(dx, dy);
}

// a0 = Sx(sy), gx = Dx(sy), gy = Dy(sx) with sy = Sy(I), sx = Sx(I): the
// adjoint first forms g_sy = Sx^T g_a0 + Dx^T g_gx and g_sx = Dy^T g_gy.
__global__ void __launch_bounds__(BX* BY)
    k_ad_image_model_bwd1(const float* ga0, const float2* gg, float* gsy, float* gsx, Dims d) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const float s1 = corr5_adjoint(x, d.W, WS{}, [&](int i) { return ga0[at(b, y, i, d)]; });
    const float s2 = corr5_adjoint(x, d.W, WD{}, [&](int i) { return gg[at(b, y, i, d)].x; });
    gsy[at(b, y, x, d)] = s1 + s2;
    gsx[at(b, y, x, d)] = corr5_adjoint(y, d.H, WD{}, [&](int i) { return gg[at(b, i, x, d)].y; });
}

// g_I = Sy^T g_sy + Sx^T g_sx.
__global__ void __launch_bounds__(BX* BY)
    k_ad_image_model_bwd2(const float* gsy, const float* gsx, float* gI, Dims d, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const float a = corr5_adjoint(y, d.H, WS{}, [&](int i) { return gsy[at(b, i, x, d)]; });
    const float c = corr5_adjoint(x, d.W, WS{}, [&](int i) { return gsx[at(b, y, i, d)]; });
    put(gI, at(b, y, x, d), a + c, acc);
}

// ----------------------------------------------------------- float pyramid

// imageDownX_float_k / imageDownY_float_k (pyramid_k.cu:81-145): out of size
// (H, W/2) from (H, W) along x, or (H/2, W) from (H, W) along y.
template <int AXIS>
__global__ void __launch_bounds__(BX* BY) k_ad_down(const float* in, float* out, Dims di, Dims dout) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= dout.W || y >= dout.H) return;
    float tm, t0, tp;
    if (AXIS == 0) {
        tm = in[at(b, y, clampi(2 * x - 1, 0, di.W - 1), di)];
        t0 = in[at(b, y, clampi(2 * x, 0, di.W - 1), di)];
        tp = in[at(b, y, clampi(2 * x + 1, 0, di.W - 1), di)];
    } else {
        tm = in[at(b, clampi(2 * y - 1, 0, di.H - 1), x, di)];
        t0 = in[at(b, clampi(2 * y, 0, di.H - 1), x, di)];
        tp = in[at(b, clampi(2 * y + 1, 0, di.H - 1), x, di)];
    }
    const float s = tm + tp;
    out[at(b, y, x, dout)] = (float)(0.5 * (double)t0 + 0.25 * (double)s);
}

// Adjoint of k_ad_down: input j receives 0.5 from the output whose centre
// tap it is and 0.25 from outputs whose side taps (clamped) read it.
template <int AXIS>
__global__ void __launch_bounds__(BX* BY) k_ad_down_bwd(const float* gout, float* gin, Dims di, Dims dout, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= di.W || y >= di.H) return;
    const int j = AXIS == 0 ? x : y, n = AXIS == 0 ? di.W : di.H, no = AXIS == 0 ? dout.W : dout.H;
    float s = 0.0f;
    for (int o = (j - 1) / 2 - 1; o <= j / 2 + 1; ++o) {
        if (o < 0 || o >= no) continue;
        const float go = AXIS == 0 ? gout[at(b, y, o, dout)] : gout[at(b, o, x, dout)];
        float w = 0.0f;
        if (clampi(2 * o, 0, n - 1) == j) w += 0.5f;
        if (clampi(2 * o - 1, 0, n - 1) == j) w += 0.25f;
        if (clampi(2 * o + 1, 0, n - 1) == j) w += 0.25f;
        s = fmaf(w, go, s);
    }
    put(gin, at(b, y, x, di), s, acc);
}

// ------------------------------------------------------------ propagation

// The fields advected by one pass at one pixel: flow, scalar and vector
// payloads.
struct Px5 {
    float2 f;
    float s;
    float2 v;
};

// One upwind update of a pixel from its neighbours m, z, p along the pass
// (propagation_k.cu:45-53 and the payload kernels): the speed is the flow
// component c (0: x, 1: y) of the neighbour with the larger magnitude;
// pixels outside the border keep z.
template <int C, bool PAY>
__device__ __forceinline__ Px5 upwind_px(const Px5& m, const Px5& z, const Px5& p, float dt, bool in) {
    const float um = C == 0 ? m.f.x : m.f.y, up = C == 0 ? p.f.x : p.f.y;
    const float Ud = (fabsf(up) - fabsf(um)) > 0.0f ? up : um;
    const float nd = -(dt * Ud);
    const bool back = Ud >= 0.0f;
    Px5 o;
    o.f = in ? make_float2(upwind(m.f.x, z.f.x, p.f.x, back, nd), upwind(m.f.y, z.f.y, p.f.y, back, nd)) : z.f;
    if constexpr (PAY) {
        o.s = in ? upwind(m.s, z.s, p.s, back, nd) : z.s;
        o.v = in ? make_float2(upwind(m.v.x, z.v.x, p.v.x, back, nd), upwind(m.v.y, z.v.y, p.v.y, back, nd)) : z.v;
    }
    return o;
}

template <bool PAY>
__device__ __forceinline__ Px5 load_px(const float2* f, const float* s, const float2* v, size_t i) {
    Px5 r;
    r.f = f[i];
    if constexpr (PAY) {
        r.s = s[i];
        r.v = v[i];
    }
    return r;
}

// Lane i of a wave receives lane i-1's value (PREV) or lane i+1's; the lane
// without a source receives 0. Every lane of the wave must be active.
template <bool PREV>
__device__ __forceinline__ float lane_shift(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), PREV ? 0x138 : 0x130, 0xf, 0xf, true));
}
template <bool PREV, bool PAY>
__device__ __forceinline__ Px5 lane_shift_px(const Px5& z) {
    Px5 r;
    r.f = make_float2(lane_shift<PREV>(z.f.x), lane_shift<PREV>(z.f.y));
    if constexpr (PAY) {
        r.s = lane_shift<PREV>(z.s);
        r.v = make_float2(lane_shift<PREV>(z.v.x), lane_shift<PREV>(z.v.y));
    }
    return r;
}

template <bool PAY>
__device__ __forceinline__ void store_px(float2* f, float* s, float2* v, size_t i, const Px5& r) {
    f[i] = r.f;
    if constexpr (PAY) {
        s[i] = r.s;
        v[i] = r.v;
    }
}

// One upwind pass of the flow and, with payloads, of a scalar and a vector
// field (flowPropagate{X,Y}_k / flowPropagatePayload{X,Y}_k,
// propagation_k.cu:16-334).
template <int AXIS, bool PAY>
__global__ void __launch_bounds__(BX* BY) k_ad_prop(const float2* fin, const float* sin, const float2* vin,
                                                    float2* fout, float* sout, float2* vout, Dims d, float dt,
                                                    int border) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const int ym = AXIS == 1 ? clampi(y - 1, 0, d.H - 1) : y, yp = AXIS == 1 ? clampi(y + 1, 0, d.H - 1) : y;
    const int xm = AXIS == 0 ? clampi(x - 1, 0, d.W - 1) : x, xp = AXIS == 0 ? clampi(x + 1, 0, d.W - 1) : x;
    const size_t i0 = at(b, y, x, d);
    const bool in = (x >= border && x < d.W - border) && (y >= border && y < d.H - border);
    const Px5 o = upwind_px<AXIS, PAY>(load_px<PAY>(fin, sin, vin, at(b, ym, xm, d)), load_px<PAY>(fin, sin, vin, i0),
                                       load_px<PAY>(fin, sin, vin, at(b, yp, xp, d)), dt, in);
    store_px<PAY>(fout, sout, vout, i0, o);
}

// An X pass followed by a Y pass in one launch: a workgroup of 64 x 16
// threads (a wave per row) computes the X pass on 16 rows (clamped into the
// image) into LDS and the Y pass on the 14 inner ones, so the pair reads and
// writes each field once instead of twice. Each lane loads its own pixel
// (lanes past the last column load that column, the clamped read) and takes
// its row neighbours from the adjacent lanes; only the first and last lane
// load theirs. With SAVE, the X pass's result (the Y pass's input, which the
// adjoint needs) is written too. Same arithmetic as two k_ad_prop launches.
constexpr int kPairRows = 14;
template <bool PAY, bool SAVE>
__global__ void __launch_bounds__(64 * (kPairRows + 2))
    k_ad_prop_pair(const float2* fin, const float* sin, const float2* vin, float2* fmid, float* smid, float2* vmid,
                   float2* fout, float* sout, float2* vout, Dims d, float dt, int border) {
    __shared__ Px5 L[kPairRows + 2][64];
    const int tx = threadIdx.x, ty = wave_y(), b = blockIdx.z;
    const int x = blockIdx.x * 64 + tx;
    const int yr = (int)blockIdx.y * kPairRows - 1 + ty;  // row of this thread's X pass
    const int yc = clampi(yr, 0, d.H - 1);
    const bool xok = x < d.W;
    const bool xin = x >= border && x < d.W - border;
    {
        const Px5 z = load_px<PAY>(fin, sin, vin, at(b, yc, clampi(x, 0, d.W - 1), d));
        Px5 m = lane_shift_px<true, PAY>(z), p = lane_shift_px<false, PAY>(z);
        if (tx == 0) m = load_px<PAY>(fin, sin, vin, at(b, yc, clampi(x - 1, 0, d.W - 1), d));
        if (tx == 63) p = load_px<PAY>(fin, sin, vin, at(b, yc, clampi(x + 1, 0, d.W - 1), d));
        const bool in = xin && yc >= border && yc < d.H - border;
        const Px5 o = upwind_px<0, PAY>(m, z, p, dt, in);
        L[ty][tx] = o;
        if (SAVE && xok && ty >= 1 && ty <= kPairRows && yr < d.H) store_px<PAY>(fmid, smid, vmid, at(b, yc, x, d), o);
    }
    __syncthreads();
    if (xok && ty >= 1 && ty <= kPairRows && yr < d.H) {
        const bool in = xin && yr >= border && yr < d.H - border;
        const Px5 o = upwind_px<1, PAY>(L[ty - 1][tx], L[ty][tx], L[ty + 1][tx], dt, in);
        store_px<PAY>(fout, sout, vout, at(b, yr, x, d), o);
    }
}

// What an output of a pass contributes to its three inputs (see out_terms).
struct OutTerms {
    float cm, cz, cp;  // coefficients of the inputs at o-1, o, o+1
    float gud;         // adjoint of the speed
    unsigned flags;    // bit 0: o inside the border; bit 1: speed taken from o+1
};

// Rows of pixels (Px5) in LDS as planes (flow, scalar, vector), 20 bytes
// per pixel with consecutive lanes at consecutive addresses (an array of
// Px5 takes 24 bytes, its padded size, and strides the banks). A row can
// hold a pixel or an output's OutTerms instead.
template <int S>
struct LdsRows {
    float2 f[S][64];
    float s[S][64];
    float2 v[S][64];
    template <bool PAY>
    __device__ __forceinline__ void put(int i, int tx, const Px5& p) {
        f[i][tx] = p.f;
        if constexpr (PAY) {
            s[i][tx] = p.s;
            v[i][tx] = p.v;
        }
    }
    template <bool PAY>
    __device__ __forceinline__ Px5 get(int i, int tx) const {
        Px5 r;
        r.f = f[i][tx];
        r.s = PAY ? s[i][tx] : 0.f;
        r.v = PAY ? v[i][tx] : make_float2(0.f, 0.f);
        return r;
    }
    __device__ __forceinline__ void put_terms(int i, int tx, const OutTerms& t) {
        f[i][tx] = make_float2(t.cm, t.cz);
        s[i][tx] = t.cp;
        v[i][tx] = make_float2(t.gud, __uint_as_float(t.flags));
    }
    __device__ __forceinline__ OutTerms get_terms(int i, int tx) const {
        const float2 a = f[i][tx], c = v[i][tx];
        OutTerms t;
        t.cm = a.x;
        t.cz = a.y;
        t.cp = s[i][tx];
        t.gud = c.x;
        t.flags = __float_as_uint(c.y);
        return t;
    }
};

// Two propagation iterations (X, Y, X, Y passes) in one launch, for the
// forward of frames whose intermediates are not kept: a workgroup of
// 64 x NW threads covers 64 columns (lane tx: column x0-1+tx) and S = NW x RPW
// staged rows (wave w: staged rows i = w + NW k, image rows y0-2+i, clamped
// into the image), and writes the S-4 inner rows of its 62 inner columns
// (lanes 1..62); it reads and writes each field once per two iterations
// instead of twice. The first X pass (row neighbours from the adjacent
// lanes, the edge lanes load theirs) goes to LDS; the first Y pass reads it
// at rows 1..S-2, the second X pass follows in registers (neighbours from
// the adjacent lanes only, so lanes 0 and 63 serve as its halo) into LDS;
// the second Y pass reads that at the output rows. A row or column outside
// the image holds the values of the edge row or column it is clamped to
// (the clamped reads of the pass that follows), and each Y pass reads its
// rows' neighbours at the clamped row indices. Every pixel's value is
// computed by the same per-pixel function from the same operands as in two
// k_ad_prop_pair launches, so the results are bitwise theirs.
constexpr int kQuadCols = 62;
// 16 waves of 2 rows: 32 staged rows for 28 output rows, two 40 KB sets of
// LDS planes (two workgroups, 32 waves, per CU). With 16 waves of one row
// (16 staged rows for 12, 24-byte Px5 in LDS, three barriers) the
// 3840x2160 launch took 94 us.
constexpr int kQuadWaves = 16, kQuadRowsPerWave = 2, kQuadRows = kQuadWaves * kQuadRowsPerWave - 4;
template <bool PAY, int NW, int RPW>
__global__ void __launch_bounds__(64 * NW)
    k_ad_prop_quad(const float2* fin, const float* sin, const float2* vin, float2* fout, float* sout, float2* vout,
                   Dims d, float dt, int border) {
    constexpr int S = NW * RPW;
    __shared__ LdsRows<S> LA;  // first X pass
    __shared__ LdsRows<S> LB;  // second X pass
    const int tx = threadIdx.x, w = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * kQuadCols - 1 + tx;  // column of this lane
    const int y0 = (int)blockIdx.y * (S - 4);
    const int xc = clampi(x, 0, d.W - 1);
    const bool xin = x >= border && x < d.W - border;
    // staged row holding image row r (r within the staged range)
    auto srow = [&](int r) { return clampi(r, 0, d.H - 1) - (y0 - 2); };
    int yr[RPW], yc[RPW];
    bool in[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        yr[k] = y0 - 2 + w + NW * k;
        yc[k] = clampi(yr[k], 0, d.H - 1);
        in[k] = xin && yc[k] >= border && yc[k] < d.H - border;
    }
    // first X pass on every staged row
    Px5 z[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) z[k] = load_px<PAY>(fin, sin, vin, at(b, yc[k], xc, d));
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        Px5 m = lane_shift_px<true, PAY>(z[k]), p = lane_shift_px<false, PAY>(z[k]);
        if (tx == 0) m = load_px<PAY>(fin, sin, vin, at(b, yc[k], clampi(x - 1, 0, d.W - 1), d));
        if (tx == 63) p = load_px<PAY>(fin, sin, vin, at(b, yc[k], clampi(x + 1, 0, d.W - 1), d));
        LA.template put<PAY>(w + NW * k, tx, upwind_px<0, PAY>(m, z[k], p, dt, in[k]));
    }
    __syncthreads();
    // first Y pass on staged rows 1..S-2 and the second X pass after it
    // (valid on lanes 1..62)
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        if (i < 1 || i > S - 2) continue;  // uniform across the wave
        const Px5 y1 = upwind_px<1, PAY>(LA.template get<PAY>(srow(yc[k] - 1), tx), LA.template get<PAY>(srow(yc[k]), tx),
                                         LA.template get<PAY>(srow(yc[k] + 1), tx), dt, in[k]);
        const Px5 m = lane_shift_px<true, PAY>(y1), p = lane_shift_px<false, PAY>(y1);
        LB.template put<PAY>(i, tx, upwind_px<0, PAY>(m, y1, p, dt, in[k]));
    }
    __syncthreads();
    // second Y pass on the S-4 output rows
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        if (i < 2 || i > S - 3 || yr[k] >= d.H) continue;
        if (tx >= 1 && tx <= kQuadCols && x < d.W) {
            const Px5 o = upwind_px<1, PAY>(LB.template get<PAY>(srow(yc[k] - 1), tx), LB.template get<PAY>(srow(yc[k]), tx),
                                            LB.template get<PAY>(srow(yc[k] + 1), tx), dt, in[k]);
            store_px<PAY>(fout, sout, vout, at(b, yr[k], x, d), o);
        }
    }
}

// Adjoint of one upwind pass. Output o = z + nd * D with nd = -dt*Ud and
// D = back ? z - m : p - z, for every field, where Ud = (|p_u| > |m_u|) ?
// p_u : m_u is read from the flow component along the pass (u). Holding
// `back` and the choice of Ud fixed:
//   d o/d z = 1 + nd*(back ? 1 : -1), d o/d m = back ? -nd : 0,
//   d o/d p = back ? 0 : nd, d o/d Ud = -dt * D (for each field),
// and Ud's derivative goes to p_u or m_u. Pixels within the border copy z.
// Input j gathers from the outputs at j-1, j, j+1 along the pass (for
// interior outputs the neighbour reads are unclamped: border >= 1).
//
// One input position j of a line of n: X holds the forward inputs at
// j-2 .. j+2 and G the output gradients at j-1 .. j+1 (values at positions
// outside the line are never used); interior(o) tells whether output o is
// inside the border. C: the flow component of the speed (0: X pass).
template <int C, bool PAY, class Interior>
__device__ __forceinline__ Px5 upwind_adjoint_px(const Px5 (&X)[5], const Px5 (&G)[3], int j, int n,
                                                 Interior interior, float dt) {
    float2 gf = make_float2(0.f, 0.f), gv = make_float2(0.f, 0.f);
    float gs = 0.f;
#pragma unroll
    for (int oi = 0; oi < 3; ++oi) {  // output o = j - 1 + oi
        const int o = j - 1 + oi;
        if (o < 0 || o >= n) continue;
        const float2 g_f = G[oi].f;
        const float g_s = PAY ? G[oi].s : 0.f;
        const float2 g_v = PAY ? G[oi].v : make_float2(0.f, 0.f);
        if (!interior(o)) {
            if (oi == 1) {  // out = z
                gf.x += g_f.x;
                gf.y += g_f.y;
                if (PAY) {
                    gs += g_s;
                    gv.x += g_v.x;
                    gv.y += g_v.y;
                }
            }
            continue;
        }
        // o - 1, o, o + 1 are window positions oi, oi + 1, oi + 2
        const float2 fm = X[oi].f, f0 = X[oi + 1].f, fp = X[oi + 2].f;
        const float um = C == 0 ? fm.x : fm.y, up = C == 0 ? fp.x : fp.y;
        const bool selp = (fabsf(up) - fabsf(um)) > 0.0f;
        const float Ud = selp ? up : um;
        const float nd = -(dt * Ud);
        const bool back = Ud >= 0.0f;
        // coefficient of this input in output o, by position of j in {m, z, p}
        const float c = (oi == 1) ? (1.0f + (back ? nd : -nd)) : ((oi == 2) ? (back ? -nd : 0.f) : (back ? 0.f : nd));
        gf.x = fmaf(c, g_f.x, gf.x);
        gf.y = fmaf(c, g_f.y, gf.y);
        // adjoint of Ud: sum over fields of g_out * (-dt * D)
        auto D = [&](float m, float z, float p) { return back ? (z - m) : (p - z); };
        float gUd = g_f.x * D(fm.x, f0.x, fp.x) + g_f.y * D(fm.y, f0.y, fp.y);
        if constexpr (PAY) {
            gs = fmaf(c, g_s, gs);
            gv.x = fmaf(c, g_v.x, gv.x);
            gv.y = fmaf(c, g_v.y, gv.y);
            gUd += g_s * D(X[oi].s, X[oi + 1].s, X[oi + 2].s) + g_v.x * D(X[oi].v.x, X[oi + 1].v.x, X[oi + 2].v.x) +
                   g_v.y * D(X[oi].v.y, X[oi + 1].v.y, X[oi + 2].v.y);
        }
        gUd *= -dt;
        // Ud is the u component of the neighbour on the chosen side
        const bool from = selp ? (oi == 0) : (oi == 2);
        if (from) {
            if (C == 0) gf.x += gUd;
            else gf.y += gUd;
        }
    }
    Px5 r;
    r.f = gf;
    r.s = gs;
    r.v = gv;
    return r;
}

// upwind_adjoint_px in two halves, for kernels that hold every output of a
// pass: out_terms computes, once per output o, what o contributes to its
// three inputs (the coefficients of the inputs at o-1, o, o+1 and the
// adjoint of its speed, routed to the neighbour the speed came from), and
// gather_px sums the terms of outputs j-1, j, j+1 for input j. Both use
// upwind_adjoint_px's expressions in its order, so together they give its
// results bit for bit, while each output's terms are computed once instead
// of once per input that reads them (OutTerms above).
template <int C, bool PAY>
__device__ __forceinline__ OutTerms out_terms(const Px5& xm, const Px5& x0, const Px5& xp, const Px5& g, bool interior,
                                              float dt) {
    OutTerms t;
    t.cm = t.cz = t.cp = t.gud = 0.f;
    t.flags = 0u;
    if (!interior) return t;
    const float um = C == 0 ? xm.f.x : xm.f.y, up = C == 0 ? xp.f.x : xp.f.y;
    const bool selp = (fabsf(up) - fabsf(um)) > 0.0f;
    const float Ud = selp ? up : um;
    const float nd = -(dt * Ud);
    const bool back = Ud >= 0.0f;
    t.cz = 1.0f + (back ? nd : -nd);
    t.cm = back ? -nd : 0.f;
    t.cp = back ? 0.f : nd;
    auto D = [&](float m, float z, float p) { return back ? (z - m) : (p - z); };
    float gUd = g.f.x * D(xm.f.x, x0.f.x, xp.f.x) + g.f.y * D(xm.f.y, x0.f.y, xp.f.y);
    if constexpr (PAY)
        gUd += g.s * D(xm.s, x0.s, xp.s) + g.v.x * D(xm.v.x, x0.v.x, xp.v.x) + g.v.y * D(xm.v.y, x0.v.y, xp.v.y);
    gUd *= -dt;
    t.gud = gUd;
    t.flags = 1u | (selp ? 2u : 0u);
    return t;
}

// Input j from outputs j-1, j, j+1 (T, G; valid: the output is inside the
// line).
template <int C, bool PAY>
__device__ __forceinline__ Px5 gather_px(const OutTerms (&T)[3], const Px5 (&G)[3], const bool (&valid)[3]) {
    float2 gf = make_float2(0.f, 0.f), gv = make_float2(0.f, 0.f);
    float gs = 0.f;
#pragma unroll
    for (int oi = 0; oi < 3; ++oi) {
        if (!valid[oi]) continue;
        const Px5& g = G[oi];
        if (!(T[oi].flags & 1u)) {
            if (oi == 1) {  // out = z
                gf.x += g.f.x;
                gf.y += g.f.y;
                if (PAY) {
                    gs += g.s;
                    gv.x += g.v.x;
                    gv.y += g.v.y;
                }
            }
            continue;
        }
        const float c = (oi == 1) ? T[oi].cz : ((oi == 2) ? T[oi].cm : T[oi].cp);
        gf.x = fmaf(c, g.f.x, gf.x);
        gf.y = fmaf(c, g.f.y, gf.y);
        if constexpr (PAY) {
            gs = fmaf(c, g.s, gs);
            gv.x = fmaf(c, g.v.x, gv.x);
            gv.y = fmaf(c, g.v.y, gv.y);
        }
        const bool selp = (T[oi].flags & 2u) != 0u;
        const bool from = selp ? (oi == 0) : (oi == 2);
        if (from) {
            if (C == 0) gf.x += T[oi].gud;
            else gf.y += T[oi].gud;
        }
    }
    Px5 r;
    r.f = gf;
    r.s = gs;
    r.v = gv;
    return r;
}

template <bool PREV>
__device__ __forceinline__ OutTerms lane_shift_terms(const OutTerms& t) {
    OutTerms r;
    r.cm = lane_shift<PREV>(t.cm);
    r.cz = lane_shift<PREV>(t.cz);
    r.cp = lane_shift<PREV>(t.cp);
    r.gud = lane_shift<PREV>(t.gud);
    r.flags = (unsigned)__builtin_amdgcn_update_dpp(0, (int)t.flags, PREV ? 0x138 : 0x130, 0xf, 0xf, true);
    return r;
}

template <bool PAY>
__device__ __forceinline__ Px5 load_grad_px(const float2* f, const float* s, const float2* v, size_t i) {
    Px5 r;
    r.f = f ? f[i] : make_float2(0.f, 0.f);
    r.s = (PAY && s) ? s[i] : 0.f;
    r.v = (PAY && v) ? v[i] : make_float2(0.f, 0.f);
    return r;
}

template <bool PAY>
__device__ __forceinline__ void put_px(float2* f, float* s, float2* v, size_t i, const Px5& r, int acc) {
    put2(f, i, r.f, acc & 1);
    if constexpr (PAY) {
        put(s, i, r.s, acc & 2);
        put2(v, i, r.v, acc & 4);
    }
}

template <int AXIS, bool PAY>
__global__ void __launch_bounds__(BX* BY)
    k_ad_prop_bwd(const float2* fin, const float* sin, const float2* vin, const float2* gfo, const float* gso,
                  const float2* gvo, float2* gfi, float* gsi, float2* gvi, Dims d, float dt, int border, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const int n = AXIS == 0 ? d.W : d.H, j = AXIS == 0 ? x : y;
    auto idx = [&](int k) { return AXIS == 0 ? at(b, y, k, d) : at(b, k, x, d); };
    auto interior = [&](int k) {
        const int xx = AXIS == 0 ? k : x, yy = AXIS == 0 ? y : k;
        return (xx >= border && xx < d.W - border) && (yy >= border && yy < d.H - border);
    };
    // everything this input can reach, loaded once up front (reads clamped
    // to the line)
    Px5 X[5], G[3];
#pragma unroll
    for (int k = 0; k < 5; ++k) X[k] = load_px<PAY>(fin, sin, vin, idx(clampi(j - 2 + k, 0, n - 1)));
#pragma unroll
    for (int k = 0; k < 3; ++k) G[k] = load_grad_px<PAY>(gfo, gso, gvo, idx(clampi(j - 1 + k, 0, n - 1)));
    put_px<PAY>(gfi, gsi, gvi, at(b, y, x, d), upwind_adjoint_px<AXIS, PAY>(X, G, j, n, interior, dt), acc);
}

// Adjoint of an X pass followed by a Y pass (the pair of k_ad_prop_pair):
// a workgroup of 64 x 16 threads computes the Y pass's adjoint at 64 columns
// (the 62 of its output plus one on each side) of 16 rows into LDS, then
// the X pass's adjoint at the 62 inner columns, reading the pair's input,
// the X pass's result (mid) and the output gradients once. Same arithmetic
// as two k_ad_prop_bwd launches.
constexpr int kPairCols = 62, kPairBwdRows = 16;
template <bool PAY>
__global__ void __launch_bounds__(64 * kPairBwdRows)
    k_ad_prop_pair_bwd(const float2* fin, const float* sin, const float2* vin, const float2* fmid, const float* smid,
                       const float2* vmid, const float2* gfo, const float* gso, const float2* gvo, float2* gfi,
                       float* gsi, float2* gvi, Dims d, float dt, int border, int acc) {
    __shared__ Px5 LG[kPairBwdRows][64];
    const int tx = threadIdx.x, ty = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * kPairCols - 1 + tx;
    const int y = (int)blockIdx.y * kPairBwdRows + ty;
    const bool xin = x >= border && x < d.W - border;
    Px5 gm;
    gm.f = make_float2(0.f, 0.f);
    gm.s = 0.f;
    gm.v = make_float2(0.f, 0.f);
    if (x >= 0 && x < d.W && y < d.H) {
        Px5 X[5], G[3];
#pragma unroll
        for (int k = 0; k < 5; ++k) X[k] = load_px<PAY>(fmid, smid, vmid, at(b, clampi(y - 2 + k, 0, d.H - 1), x, d));
#pragma unroll
        for (int k = 0; k < 3; ++k) G[k] = load_grad_px<PAY>(gfo, gso, gvo, at(b, clampi(y - 1 + k, 0, d.H - 1), x, d));
        auto interior = [&](int o) { return xin && o >= border && o < d.H - border; };
        gm = upwind_adjoint_px<1, PAY>(X, G, y, d.H, interior, dt);
    }
    LG[ty][tx] = gm;
    __syncthreads();
    if (tx >= 1 && tx <= kPairCols && x < d.W && y < d.H) {
        Px5 X[5];
        const Px5 G[3] = {LG[ty][tx - 1], LG[ty][tx], LG[ty][tx + 1]};
#pragma unroll
        for (int k = 0; k < 5; ++k) X[k] = load_px<PAY>(fin, sin, vin, at(b, y, clampi(x - 2 + k, 0, d.W - 1), d));
        const bool yin = y >= border && y < d.H - border;
        auto interior = [&](int o) { return yin && o >= border && o < d.W - border; };
        put_px<PAY>(gfi, gsi, gvi, at(b, y, x, d), upwind_adjoint_px<0, PAY>(X, G, x, d.W, interior, dt), acc);
    }
}

// The same adjoint without the X pass's result. A workgroup of 64 x 16
// threads (a wave per row) covers 64 columns: lane tx is column x0-2+tx. The
// pair's input is loaded once per row of this wave (the own row, y) and of
// two rows above and below the workgroup's 16 (the first four waves), with
// reads clamped into the image as in the forward; the X pass is rerun on
// those 20 rows with the row neighbours taken from the adjacent lanes (valid
// on lanes 1..62) into LDS, with the output gradients of rows y0-1 ..
// y0+16. Then the Y pass's adjoint at lanes 1..62, and the X pass's at lanes
// 2..61 (columns x0 .. x0+59) with its inputs and the Y adjoint of the
// neighbouring columns again taken from adjacent lanes. Every value comes
// from the same per-pixel functions and operands as in the forward and in
// k_ad_prop_pair_bwd, so the results are bitwise theirs; the X pass's
// result (20 bytes per pixel written forward and read back) is not needed.
constexpr int kPairRcCols = 60, kPairRcRows = kPairBwdRows + 4;
template <bool PAY>
__global__ void __launch_bounds__(64 * kPairBwdRows)
    k_ad_prop_pair_bwd_rc(const float2* fin, const float* sin, const float2* vin, const float2* gfo,
                          const float* gso, const float2* gvo, float2* gfi, float* gsi, float2* gvi, Dims d, float dt,
                          int border, int acc) {
    __shared__ Px5 LM[kPairRcRows][64];        // X pass of rows y0-2 .. y0+17
    __shared__ Px5 LG[kPairBwdRows + 2][64];   // output gradients of rows y0-1 .. y0+16
    const int tx = threadIdx.x, ty = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * kPairRcCols - 2 + tx;  // column of this lane
    const int y0 = (int)blockIdx.y * kPairBwdRows, y = y0 + ty;
    const int xc = clampi(x, 0, d.W - 1);
    const bool xin = x >= border && x < d.W - border;
    // X pass of staged row i (image row clamp(y0-2+i)); returns its input
    auto x_pass = [&](int i) {
        const int yc = clampi(y0 - 2 + i, 0, d.H - 1);
        const Px5 z = load_px<PAY>(fin, sin, vin, at(b, yc, xc, d));
        const Px5 m = lane_shift_px<true, PAY>(z), p = lane_shift_px<false, PAY>(z);
        const bool in = xin && yc >= border && yc < d.H - border;
        LM[i][tx] = upwind_px<0, PAY>(m, z, p, dt, in);
        return z;
    };
    const Px5 zown = x_pass(ty + 2);  // the pair's input at (y, x)
    if (ty < 4) x_pass(ty < 2 ? ty : kPairBwdRows + ty);  // rows 0, 1, 18, 19
    LG[ty + 1][tx] = load_grad_px<PAY>(gfo, gso, gvo, at(b, clampi(y, 0, d.H - 1), xc, d));
    if (ty == 0) LG[0][tx] = load_grad_px<PAY>(gfo, gso, gvo, at(b, clampi(y0 - 1, 0, d.H - 1), xc, d));
    if (ty == kPairBwdRows - 1)
        LG[kPairBwdRows + 1][tx] = load_grad_px<PAY>(gfo, gso, gvo, at(b, clampi(y0 + kPairBwdRows, 0, d.H - 1), xc, d));
    __syncthreads();
    // adjoint of the Y pass (zero outside the image)
    Px5 gm;
    gm.f = make_float2(0.f, 0.f);
    gm.s = 0.f;
    gm.v = make_float2(0.f, 0.f);
    if (tx >= 1 && tx <= 62 && x >= 0 && x < d.W && y < d.H) {
        Px5 X[5], G[3];
#pragma unroll
        for (int k = 0; k < 5; ++k) X[k] = LM[ty + k][tx];  // rows clamp(y-2+k)
#pragma unroll
        for (int k = 0; k < 3; ++k) G[k] = LG[ty + k][tx];  // rows clamp(y-1+k)
        auto interior = [&](int o) { return xin && o >= border && o < d.H - border; };
        gm = upwind_adjoint_px<1, PAY>(X, G, y, d.H, interior, dt);
    }
    // adjoint of the X pass: the Y adjoint at columns x-1 .. x+1 and the
    // pair's input at columns clamp(x-2) .. clamp(x+2) of row y, from lanes
    // tx-1 .. tx+1 and tx-2 .. tx+2 (every lane active here)
    const Px5 G3[3] = {lane_shift_px<true, PAY>(gm), gm, lane_shift_px<false, PAY>(gm)};
    const Px5 m1 = lane_shift_px<true, PAY>(zown), p1 = lane_shift_px<false, PAY>(zown);
    const Px5 X5[5] = {lane_shift_px<true, PAY>(m1), m1, zown, p1, lane_shift_px<false, PAY>(p1)};
    if (tx >= 2 && tx < kPairRcCols + 2 && x < d.W && y < d.H) {
        const bool yin = y >= border && y < d.H - border;
        auto interior = [&](int o) { return yin && o >= border && o < d.W - border; };
        put_px<PAY>(gfi, gsi, gvi, at(b, y, x, d), upwind_adjoint_px<0, PAY>(X5, G3, x, d.W, interior, dt), acc);
    }
}


// Adjoint of k_ad_prop_quad (two iterations), recomputing the forward
// passes inside the launch. A workgroup of 64 x NW threads covers 64
// columns, lane tx being column x0-3+tx, and S = NW x RPW staged rows, wave
// w holding staged rows i = w + NW k, image rows y0-4+i (reads clamped into
// the image); R = S - 8 output rows. Forward: the first X pass on the S rows
// (row neighbours from the adjacent lanes, the edge lanes load theirs) into
// LDS, the first Y pass on rows 1..S-2, the second X pass on those rows
// (adjacent lanes; valid on lanes 1..62) into LDS. Adjoints, each as
// per-output terms (out_terms) and a gather (gather_px), which give the
// single passes' adjoints bit for bit:
//   second Y pass: terms of rows 2..S-3 (lanes 1..62; the output gradients
//     read from memory) into LDS in place of the second X pass, gathered at
//     rows 3..S-4;
//   second X pass: terms at every lane (its inputs and neighbours from the
//     adjacent lanes), gathered from the adjacent lanes at lanes 2..61;
//     the result, the first Y pass's output gradient, into LDS;
//   first Y pass: terms of rows 3..S-4 into LDS in place of the first X
//     pass, gathered at rows 4..S-5;
//   first X pass: as the second, written for lanes 3..60 (columns
//     x0 .. x0+57, rows y0 .. y0+R-1).
constexpr int kQuadBwdCols = 58;
// Workgroup shape of the launches: 16 waves of 2 rows, 32 staged rows for
// 24 output rows, two 40 KB sets of LDS planes; two workgroups per CU, so
// the kernel is held to 64 VGPRs (8 waves per SIMD). With 16 waves of one
// row (16 staged rows for 8 output rows, the gradients staged in LDS too,
// each input recomputing the terms of the three outputs that read it) the
// 3840x2160 launch took 249 us; with 8 waves of 4 rows (99 VGPRs, 4 waves
// per SIMD) 253 us.
constexpr int kQuadBwdWaves = 16, kQuadBwdRowsPerWave = 2;
template <int NW, int RPW>
struct QuadBwdShape {
    static constexpr int stage = NW * RPW, rows = stage - 8;
};
template <bool PAY, int NW, int RPW>
__global__ void __launch_bounds__(64 * NW, NW / 2)
    k_ad_prop_quad_bwd(const float2* fin, const float* sin, const float2* vin, const float2* gfo, const float* gso,
                       const float2* gvo, float2* gfi, float* gsi, float2* gvi, Dims d, float dt, int border,
                       int acc) {
    constexpr int S = QuadBwdShape<NW, RPW>::stage, R = QuadBwdShape<NW, RPW>::rows;
    __shared__ LdsRows<S> LA;  // first X pass; later the first Y pass's terms
    __shared__ LdsRows<S> LB;  // second X pass; later the second Y pass's terms, then the first Y pass's output gradient
    const int tx = threadIdx.x, w = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * kQuadBwdCols - 3 + tx;  // column of this lane
    const int y0 = (int)blockIdx.y * R;
    const int xc = clampi(x, 0, d.W - 1);
    const bool xin = x >= border && x < d.W - border;
    auto srow = [&](int r) { return clampi(r, 0, d.H - 1) - (y0 - 4); };  // staged row of image row clamp(r)
    Px5 zero;
    zero.f = make_float2(0.f, 0.f);
    zero.s = 0.f;
    zero.v = make_float2(0.f, 0.f);
    const bool xvalid[3] = {x - 1 >= 0 && x - 1 < d.W, x >= 0 && x < d.W, x + 1 >= 0 && x + 1 < d.W};
    // this wave's staged rows i = w + NW k, image rows y0-4+i
    Px5 F[RPW], Y1[RPW], gy1[RPW];
    int yr[RPW], yc[RPW];
    bool in[RPW], pix[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        yr[k] = y0 - 4 + i;
        yc[k] = clampi(yr[k], 0, d.H - 1);
        in[k] = xin && yc[k] >= border && yc[k] < d.H - border;
        pix[k] = x >= 0 && x < d.W && yr[k] >= 0 && yr[k] < d.H;
    }
    auto yinterior = [&](int o) { return xin && o >= border && o < d.H - border; };

    // ---- forward
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        F[k] = load_px<PAY>(fin, sin, vin, at(b, yc[k], xc, d));
        Px5 m = lane_shift_px<true, PAY>(F[k]), p = lane_shift_px<false, PAY>(F[k]);
        if (tx == 0) m = load_px<PAY>(fin, sin, vin, at(b, yc[k], clampi(x - 1, 0, d.W - 1), d));
        if (tx == 63) p = load_px<PAY>(fin, sin, vin, at(b, yc[k], clampi(x + 1, 0, d.W - 1), d));
        LA.template put<PAY>(i, tx, upwind_px<0, PAY>(m, F[k], p, dt, in[k]));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        const bool mid = i >= 1 && i <= S - 2;  // rows of the first Y and second X pass
        Y1[k] = zero;
        if (mid)
            Y1[k] = upwind_px<1, PAY>(LA.template get<PAY>(srow(yc[k] - 1), tx), LA.template get<PAY>(srow(yc[k]), tx),
                                      LA.template get<PAY>(srow(yc[k] + 1), tx), dt, in[k]);
        const Px5 m = lane_shift_px<true, PAY>(Y1[k]), p = lane_shift_px<false, PAY>(Y1[k]);
        if (mid) LB.template put<PAY>(i, tx, upwind_px<0, PAY>(m, Y1[k], p, dt, in[k]));
    }
    __syncthreads();

    // ---- second Y pass: terms of output rows 2 .. S-3 (inside the image)
    OutTerms t[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        t[k] = out_terms<1, PAY>(zero, zero, zero, zero, false, dt);
        if (i >= 2 && i <= S - 3 && pix[k] && tx >= 1 && tx <= 62) {
            const Px5 g = load_grad_px<PAY>(gfo, gso, gvo, at(b, yr[k], xc, d));
            t[k] = out_terms<1, PAY>(LB.template get<PAY>(srow(yr[k] - 1), tx), LB.template get<PAY>(srow(yr[k]), tx),
                                     LB.template get<PAY>(srow(yr[k] + 1), tx), g, yinterior(yr[k]), dt);
        }
    }
    __syncthreads();  // every read of LB (second X pass) is done
#pragma unroll
    for (int k = 0; k < RPW; ++k) LB.put_terms(w + NW * k, tx, t[k]);
    __syncthreads();

    // ---- gathers of the second pair (rows 3 .. S-4)
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        const bool rows2 = i >= 3 && i <= S - 4;
        Px5 g_x2 = zero;
        if (rows2 && pix[k] && tx >= 1 && tx <= 62) {
            OutTerms T[3];
            Px5 G[3];
            bool valid[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int o = yr[k] - 1 + q;  // output row, staged row i-1+q
                valid[q] = o >= 0 && o < d.H;
                T[q] = LB.get_terms(i - 1 + q, tx);
                G[q] = load_grad_px<PAY>(gfo, gso, gvo, at(b, clampi(o, 0, d.H - 1), xc, d));
            }
            g_x2 = gather_px<1, PAY>(T, G, valid);
        }
        // second X pass: terms at this lane's column, gathered from the
        // adjacent lanes
        const bool yin = yr[k] >= border && yr[k] < d.H - border;
        const Px5 m1 = lane_shift_px<true, PAY>(Y1[k]), p1 = lane_shift_px<false, PAY>(Y1[k]);
        const OutTerms to = out_terms<0, PAY>(m1, Y1[k], p1, g_x2, yin && x >= border && x < d.W - border, dt);
        const OutTerms T3[3] = {lane_shift_terms<true>(to), to, lane_shift_terms<false>(to)};
        const Px5 G3[3] = {lane_shift_px<true, PAY>(g_x2), g_x2, lane_shift_px<false, PAY>(g_x2)};
        gy1[k] = zero;
        if (rows2 && pix[k] && tx >= 2 && tx <= 61) gy1[k] = gather_px<0, PAY>(T3, G3, xvalid);
    }
    __syncthreads();  // every read of LB (the second Y pass's terms) is done
#pragma unroll
    for (int k = 0; k < RPW; ++k) LB.template put<PAY>(w + NW * k, tx, gy1[k]);
    __syncthreads();

    // ---- first Y pass: terms of output rows 3 .. S-4
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        t[k] = out_terms<1, PAY>(zero, zero, zero, zero, false, dt);
        if (i >= 3 && i <= S - 4 && pix[k] && tx >= 2 && tx <= 61)
            t[k] = out_terms<1, PAY>(LA.template get<PAY>(srow(yr[k] - 1), tx), LA.template get<PAY>(srow(yr[k]), tx),
                                     LA.template get<PAY>(srow(yr[k] + 1), tx), LB.template get<PAY>(i, tx), yinterior(yr[k]), dt);
    }
    __syncthreads();  // every read of LA (first X pass) is done
#pragma unroll
    for (int k = 0; k < RPW; ++k) LA.put_terms(w + NW * k, tx, t[k]);
    __syncthreads();

    // ---- gathers of the first pair (rows 4 .. S-5), written for lanes 3..60
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int i = w + NW * k;
        const bool rows1 = i >= 4 && i <= S - 5;
        Px5 g_x1 = zero;
        if (rows1 && pix[k] && tx >= 2 && tx <= 61) {
            OutTerms T[3];
            Px5 G[3];
            bool valid[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int o = yr[k] - 1 + q;
                valid[q] = o >= 0 && o < d.H;
                T[q] = LA.get_terms(i - 1 + q, tx);
                G[q] = LB.template get<PAY>(i - 1 + q, tx);
            }
            g_x1 = gather_px<1, PAY>(T, G, valid);
        }
        const bool yin = yr[k] >= border && yr[k] < d.H - border;
        const Px5 m1 = lane_shift_px<true, PAY>(F[k]), p1 = lane_shift_px<false, PAY>(F[k]);
        const OutTerms to = out_terms<0, PAY>(m1, F[k], p1, g_x1, yin && x >= border && x < d.W - border, dt);
        const OutTerms T3[3] = {lane_shift_terms<true>(to), to, lane_shift_terms<false>(to)};
        const Px5 G3[3] = {lane_shift_px<true, PAY>(g_x1), g_x1, lane_shift_px<false, PAY>(g_x1)};
        if (rows1 && pix[k] && tx >= 3 && tx < kQuadBwdCols + 3)
            put_px<PAY>(gfi, gsi, gvi, at(b, yr[k], x, d), gather_px<0, PAY>(T3, G3, xvalid), acc);
    }
}

// ----------------------------------------------------------------- update

// Adjoint of the closed-form solve (update_k.cu:43-63). The solve is
// o = A^-1 q with A = g g^T + gamma I (g = (a, b)) and q = g Y + gamma f,
// which simplifies (Sherman-Morrison) to
//   o = f + c g,   c = (Y - g.f) / (gamma + |g|^2).
// With k = 1 / (gamma + |g|^2) and h = (g_o . g) k, its vector-Jacobian
// product is
//   g_g = c g_o - h (f + 2 c g),   g_Y = h,   g_f = g_o - h g,
//   g_gamma = -c h.
// No term carries the 1/gamma of A^-1, so nothing cancels when gamma is
// much smaller than |g|^2 (the usual case: gamma is scaled by 1/255^2).
struct SolveGrad {
    float a, b, Y, fx, fy, gamma;
};
__device__ __forceinline__ SolveGrad solve2x2_adjoint(float a, float b, float Y, float fx, float fy, float gamma,
                                                      float gox, float goy) {
    const float k = 1.0f / ((gamma + a * a) + b * b);
    const float c = (Y - fmaf(a, fx, b * fy)) * k;
    const float h = fmaf(gox, a, goy * b) * k;
    SolveGrad g;
    g.a = fmaf(c, gox, -h * fmaf(2.0f * c, a, fx));
    g.b = fmaf(c, goy, -h * fmaf(2.0f * c, b, fy));
    g.Y = h;
    g.fx = fmaf(-h, a, gox);
    g.fy = fmaf(-h, b, goy);
    g.gamma = -c * h;
    return g;
}

__device__ __forceinline__ bool inside(float v, float lim) { return v > -lim && v < lim; }

// flowUpdate_k (update_k.cu:15-79) on a batch.
__global__ void __launch_bounds__(BX* BY) k_ad_update(const float* a0, const float2* g, const float* aold,
                                                      const float2* f, float2* fout, Dims d, float gamma,
                                                      float maxflow) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const size_t i = at(b, y, x, d);
    const float2 gg = g[i], ff = f[i];
    float ox, oy;
    solve2x2(gg.x, gg.y, aold[i] - a0[i], ff.x, ff.y, gamma, ox, oy);
    ox = fmaxf(-maxflow, fminf(ox, maxflow));
    oy = fmaxf(-maxflow, fminf(oy, maxflow));
    ox = not_finite(ox) ? 0.0f : ox;
    oy = not_finite(oy) ? 0.0f : oy;
    fout[i] = make_float2(ox, oy);
}

// Block sum of v into partial[blockid] (one value per workgroup).
__device__ __forceinline__ void block_sum(float v, float* partial) {
    __shared__ float red[BX * BY];
    const int t = threadIdx.y * BX + threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = (BX * BY) / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) partial[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(BX* BY)
    k_ad_update_bwd(const float* a0, const float2* g, const float* aold, const float2* f, const float2* gout,
                    float* ga0, float2* gg, float* gaold, float2* gf, float* gamma_partial, Dims d, float gamma,
                    float maxflow, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    float ggam = 0.0f;
    if (x < d.W && y < d.H) {
        const size_t i = at(b, y, x, d);
        const float2 gr = g[i], ff = f[i];
        const float Y = aold[i] - a0[i];
        float ox, oy;
        solve2x2(gr.x, gr.y, Y, ff.x, ff.y, gamma, ox, oy);
        // clamp, then sanitize: the derivative passes where the value was
        // finite and strictly inside the clamp
        const float2 go = gout[i];
        const float gox = (inside(ox, maxflow) && !not_finite(ox)) ? go.x : 0.0f;
        const float goy = (inside(oy, maxflow) && !not_finite(oy)) ? go.y : 0.0f;
        const SolveGrad s = solve2x2_adjoint(gr.x, gr.y, Y, ff.x, ff.y, gamma, gox, goy);
        put(ga0, i, -s.Y, acc & 1);
        put2(gg, i, make_float2(s.a, s.b), acc & 2);
        put(gaold, i, s.Y, acc & 4);
        put2(gf, i, make_float2(s.fx, s.fy), acc & 8);
        ggam = s.gamma;
    }
    block_sum(ggam, gamma_partial);
}

// deltaFlowUpdate_k (update_k.cu:82-173) on a batch: writes the new delta
// flow and the level's (unsmoothed) flow.
__global__ void __launch_bounds__(BX* BY)
    k_ad_delta_update(const float* a0, const float2* g, const float* sprop, const float2* dprop, const float2* upper,
                      const int* xmap, const int* ymap, float2* dout, float2* fout, Dims d, Dims du, float gamma,
                      float maxflow) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const size_t i = at(b, y, x, d);
    const float2 gg = g[i], dd = dprop[i];
    float ox, oy;
    solve2x2(gg.x, gg.y, sprop[i] - a0[i], dd.x, dd.y, gamma, ox, oy);
    ox = not_finite(ox) ? 0.0f : ox;
    oy = not_finite(oy) ? 0.0f : oy;
    const float half = 0.5f * maxflow;
    ox = fmaxf(-half, fminf(ox, half));
    oy = fmaxf(-half, fminf(oy, half));
    const float2 u = upper[at(b, ymap[y], xmap[x], du)];
    float fx = ox + 2.0f * u.x, fy = oy + 2.0f * u.y;
    fx = fmaxf(-maxflow, fminf(fx, maxflow));
    fy = fmaxf(-maxflow, fminf(fy, maxflow));
    dout[i] = make_float2(ox, oy);
    fout[i] = make_float2(fx, fy);
}

// Adjoint of the delta update for the gradients of its new delta flow
// (gdout) and flow (gfout). The upper-flow gradient is left per pixel of
// this level (gup, 2 * the passed flow gradient); k_ad_upper_gather sums it
// onto the upper level.
__global__ void __launch_bounds__(BX* BY)
    k_ad_delta_update_bwd(const float* a0, const float2* g, const float* sprop, const float2* dprop,
                          const float2* upper, const int* xmap, const int* ymap, const float2* gdout,
                          const float2* gfout, float* ga0, float2* gg, float* gsprop, float2* gdprop, float2* gup,
                          float* gamma_partial, Dims d, Dims du, float gamma, float maxflow, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    float ggam = 0.0f;
    if (x < d.W && y < d.H) {
        const size_t i = at(b, y, x, d);
        const float2 gr = g[i], dd = dprop[i];
        const float Y = sprop[i] - a0[i];
        float sx, sy;
        solve2x2(gr.x, gr.y, Y, dd.x, dd.y, gamma, sx, sy);
        const bool fx_ok = !not_finite(sx), fy_ok = !not_finite(sy);
        const float zx = fx_ok ? sx : 0.0f, zy = fy_ok ? sy : 0.0f;
        const float half = 0.5f * maxflow;
        const float ox = fmaxf(-half, fminf(zx, half)), oy = fmaxf(-half, fminf(zy, half));
        const float2 u = upper[at(b, ymap[y], xmap[x], du)];
        const float px_ = ox + 2.0f * u.x, py_ = oy + 2.0f * u.y;
        const float2 gfo = gfout[i], gdo = gdout ? gdout[i] : make_float2(0.f, 0.f);
        // through the final clamp
        const float gpx = inside(px_, maxflow) ? gfo.x : 0.0f, gpy = inside(py_, maxflow) ? gfo.y : 0.0f;
        put2(gup, i, make_float2(2.0f * gpx, 2.0f * gpy), false);
        // o = clamp(z, half) feeds both outputs
        const float gox = gpx + gdo.x, goy = gpy + gdo.y;
        const float gzx = (fx_ok && inside(zx, half)) ? gox : 0.0f;
        const float gzy = (fy_ok && inside(zy, half)) ? goy : 0.0f;
        const SolveGrad s = solve2x2_adjoint(gr.x, gr.y, Y, dd.x, dd.y, gamma, gzx, gzy);
        put(ga0, i, -s.Y, acc & 1);
        put2(gg, i, make_float2(s.a, s.b), acc & 2);
        put(gsprop, i, s.Y, acc & 4);
        put2(gdprop, i, make_float2(s.fx, s.fy), acc & 8);
        ggam = s.gamma;
    }
    block_sum(ggam, gamma_partial);
}

// Upper-level gradient: upper pixel (X, Y) sums the per-pixel gradients of
// the pixels (x, y) of this level with xmap[x] == X and ymap[y] == Y; the
// maps are non-decreasing, so those are the ranges [xlo[X], xhi[X]) and
// [ylo[Y], yhi[Y]).
__global__ void __launch_bounds__(BX* BY)
    k_ad_upper_gather(const float2* gup, const int* xlo, const int* xhi, const int* ylo, const int* yhi, float2* gupper,
                      Dims d, Dims du, int acc) {
    const int X = blockIdx.x * BX + threadIdx.x, Y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (X >= du.W || Y >= du.H) return;
    float2 s = make_float2(0.f, 0.f);
    for (int y = ylo[Y]; y < yhi[Y]; ++y)
        for (int x = xlo[X]; x < xhi[X]; ++x) {
            const float2 v = gup[at(b, y, x, d)];
            s.x += v.x;
            s.y += v.y;
        }
    put2(gupper, at(b, Y, X, du), s, acc);
}

// -------------------------------------------------------------- smoothing

// flowSmoothX_k / flowSmoothY_k (flowsmoothing_k.cu:18-87) on a batch.
template <int AXIS>
__global__ void __launch_bounds__(BX* BY) k_ad_smooth(const float2* in, float2* out, Dims d) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int c = -2; c <= 2; ++c) {
        const float2 v = AXIS == 0 ? in[at(b, y, clampi(x + c, 0, d.W - 1), d)] : in[at(b, clampi(y + c, 0, d.H - 1), x, d)];
        sx = fmaf(kBox, v.x, sx);
        sy = fmaf(kBox, v.y, sy);
    }
    out[at(b, y, x, d)] = make_float2(sx, sy);
}

// Adjoint of one smoothing pass of the sum gout + gout2 (gout2 may be null).
template <int AXIS>
__global__ void __launch_bounds__(BX* BY)
    k_ad_smooth_bwd(const float2* gout, const float2* gout2, float2* gin, Dims d, int acc) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), b = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    auto g = [&](size_t i) {
        float2 v = gout[i];
        if (gout2) {
            const float2 u = gout2[i];
            v.x += u.x;
            v.y += u.y;
        }
        return v;
    };
    float2 s;
    if (AXIS == 0) {
        s.x = corr5_adjoint(x, d.W, WB{}, [&](int i) { return g(at(b, y, i, d)).x; });
        s.y = corr5_adjoint(x, d.W, WB{}, [&](int i) { return g(at(b, y, i, d)).y; });
    } else {
        s.x = corr5_adjoint(y, d.H, WB{}, [&](int i) { return g(at(b, i, x, d)).x; });
        s.y = corr5_adjoint(y, d.H, WB{}, [&](int i) { return g(at(b, i, x, d)).y; });
    }
    put2(gin, at(b, y, x, d), s, acc & 1);
}

// One smoothing iteration, an X and then a Y pass, in one launch: a
// workgroup of 64 x 16 threads computes the X pass on 16 rows (clamped into
// the image) into LDS and the Y pass on the inner 12, reading and writing
// the flow once instead of twice. Same arithmetic as two k_ad_smooth
// launches. (Taking the x taps from the adjacent lanes instead of loading
// them was measured slower here: the five float2 loads are served by the
// caches.)
constexpr int kSmoothRows = 12;
__global__ void __launch_bounds__(64 * (kSmoothRows + 4)) k_ad_smooth_pair(const float2* in, float2* out, Dims d) {
    __shared__ float2 L[kSmoothRows + 4][64];
    const int tx = threadIdx.x, ty = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * 64 + tx;
    const int yr = (int)blockIdx.y * kSmoothRows - 2 + ty;  // row of this thread's X pass
    const int yc = clampi(yr, 0, d.H - 1);
    if (x < d.W) {
        float sx = 0.0f, sy = 0.0f;
#pragma unroll
        for (int c = -2; c <= 2; ++c) {
            const float2 v = in[at(b, yc, clampi(x + c, 0, d.W - 1), d)];
            sx = fmaf(kBox, v.x, sx);
            sy = fmaf(kBox, v.y, sy);
        }
        L[ty][tx] = make_float2(sx, sy);
    }
    __syncthreads();
    // row yr reads rows clamp(yr + c): staged rows ty + c
    if (x < d.W && ty >= 2 && ty < kSmoothRows + 2 && yr < d.H) {
        float sx = 0.0f, sy = 0.0f;
#pragma unroll
        for (int c = -2; c <= 2; ++c) {
            const float2 v = L[ty + c][tx];
            sx = fmaf(kBox, v.x, sx);
            sy = fmaf(kBox, v.y, sy);
        }
        out[at(b, yr, x, d)] = make_float2(sx, sy);
    }
}

// Adjoint of k_ad_smooth_pair for the output gradient gout + gout2 (gout2
// may be null): a workgroup of 64 x 16 threads stages the summed output
// gradient of its rows and two on each side in LDS (each element loaded
// once), takes the Y pass's adjoint at 64 columns (its 60 output columns and
// two on each side) of 16 rows into LDS, then the X pass's adjoint at the
// inner 60. Same arithmetic as k_ad_smooth_bwd<1> followed by
// k_ad_smooth_bwd<0>.
constexpr int kSmoothBwdCols = 60, kSmoothBwdRows = 16;
__global__ void __launch_bounds__(64 * kSmoothBwdRows)
    k_ad_smooth_pair_bwd(const float2* gout, const float2* gout2, float2* gin, Dims d, int acc) {
    __shared__ float2 LGo[kSmoothBwdRows + 4][64];  // gout + gout2 of rows y0-2 .. y0+17
    __shared__ float2 L[kSmoothBwdRows][64];
    const int tx = threadIdx.x, ty = wave_y(), b = blockIdx.z;
    const int x = (int)blockIdx.x * kSmoothBwdCols - 2 + tx;
    const int y0 = (int)blockIdx.y * kSmoothBwdRows, y = y0 + ty;
    const int xc = clampi(x, 0, d.W - 1);
    // each row of output gradients loaded once (clamped; rows and columns
    // outside the image are never read below)
    auto stage = [&](int r) {
        const size_t i = at(b, clampi(y0 - 2 + r, 0, d.H - 1), xc, d);
        float2 v = gout[i];
        if (gout2) {
            const float2 u = gout2[i];
            v.x += u.x;
            v.y += u.y;
        }
        LGo[r][tx] = v;
    };
    stage(ty + 2);
    if (ty < 4) stage(ty < 2 ? ty : kSmoothBwdRows + ty);  // rows 0, 1, 18, 19
    __syncthreads();
    float2 s = make_float2(0.f, 0.f);
    if (x >= 0 && x < d.W && y < d.H) {
        // image row i is staged row i - y0 + 2
        s.x = corr5_adjoint(y, d.H, WB{}, [&](int i) { return LGo[i - y0 + 2][tx].x; });
        s.y = corr5_adjoint(y, d.H, WB{}, [&](int i) { return LGo[i - y0 + 2][tx].y; });
    }
    L[ty][tx] = s;
    __syncthreads();
    // column i of the image is lane tx + (i - x); the adjoint reads
    // columns x-2 .. x+2 inside the image only
    if (tx >= 2 && tx < kSmoothBwdCols + 2 && x < d.W && y < d.H) {
        float2 r;
        r.x = corr5_adjoint(x, d.W, WB{}, [&](int i) { return L[ty][tx + (i - x)].x; });
        r.y = corr5_adjoint(x, d.W, WB{}, [&](int i) { return L[ty][tx + (i - x)].y; });
        put2(gin, at(b, y, x, d), r, acc & 1);
    }
}

// ------------------------------------------------------------ photometric

// The photometric term of the differentiable window's loss: frame t-1
// sampled bilinearly at x - flow_t(x), the coordinates clamped into the image
// (grid_sample with align_corners and border padding), against frame t.
// Frames [B][T][H][W] float, flows [B][T][H][W] float2; z = b (T-1) + t - 1
// indexes the frame pairs.
struct Warp {
    int x0, x1, y0, y1;  // sampled columns and rows (x1 = min(x0 + 1, W - 1))
    float wx, wy;        // weights of x1, y1
    bool fx, fy;         // the coordinate is not clamped (its derivative passes)
};
__device__ __forceinline__ Warp warp_at(int x, int y, float2 f, const Dims& d) {
    Warp w;
    const float ux = (float)x - f.x, uy = (float)y - f.y;
    const float hx = (float)(d.W - 1), hy = (float)(d.H - 1);
    w.fx = !(ux < 0.f) && !(ux > hx);
    w.fy = !(uy < 0.f) && !(uy > hy);
    const float sx = ux > hx ? hx : (ux < 0.f ? 0.f : ux);
    const float sy = uy > hy ? hy : (uy < 0.f ? 0.f : uy);
    const float x0 = floorf(sx), y0 = floorf(sy);
    w.wx = sx - x0;
    w.wy = sy - y0;
    // (a NaN coordinate converts to some index: kept inside the image)
    w.x0 = clampi((int)x0, 0, d.W - 1);
    w.y0 = clampi((int)y0, 0, d.H - 1);
    w.x1 = min(w.x0 + 1, d.W - 1);
    w.y1 = min(w.y0 + 1, d.H - 1);
    return w;
}
struct Corners {
    float p00, p01, p10, p11, top, bot, v;
};
__device__ __forceinline__ Corners sample(const float* prev, const Warp& w, int W) {
    Corners c;
    c.p00 = prev[(size_t)w.y0 * W + w.x0];
    c.p01 = prev[(size_t)w.y0 * W + w.x1];
    c.p10 = prev[(size_t)w.y1 * W + w.x0];
    c.p11 = prev[(size_t)w.y1 * W + w.x1];
    c.top = (1.0f - w.wx) * c.p00 + w.wx * c.p01;
    c.bot = (1.0f - w.wx) * c.p10 + w.wx * c.p11;
    c.v = (1.0f - w.wy) * c.top + w.wy * c.bot;
    return c;
}

// Sum of the squared differences, one partial sum per workgroup.
__global__ void __launch_bounds__(BX* BY)
    k_ad_photometric(const float* frames, const float2* flows, float* partial, int T, Dims d) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), z = blockIdx.z;
    const int b = z / (T - 1), t = z - b * (T - 1) + 1;
    float e = 0.0f;
    if (x < d.W && y < d.H) {
        const size_t cur = (size_t)(b * T + t) * d.H * d.W, i = (size_t)y * d.W + x;
        const Warp w = warp_at(x, y, flows[cur + i], d);
        const Corners c = sample(frames + cur - (size_t)d.H * d.W, w, d.W);
        const float diff = frames[cur + i] - c.v;
        e = diff * diff;
    }
    block_sum(e, partial);
}

// Lane i receives lane i-1's value (PREV) or lane i+1's; 0 where there is
// none (every lane of the wave active).
template <bool PREV>
__device__ __forceinline__ int lane_shift_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, PREV ? 0x138 : 0x130, 0xf, 0xf, true);
}

// Its adjoint for the loss scale *scale (d loss / d sum of squares), for
// the frame pairs t = 1 + parity, 3 + parity, ... (z = b * npairs + k, t =
// 1 + parity + 2k): frame t gets s, the four sampled pixels of frame t-1 get
// -s times their weights, and the flow gets s times the derivative of the
// sample along each unclamped coordinate, with s = 2 *scale * difference;
// all added into gframes and gflows. Within one launch a frame is either a
// pair's target or the frame a pair samples, never both, so the target
// term is a plain add; the sampled pixels take atomic adds (neighbouring
// pixels sample the same ones), after each lane has taken over the
// contribution of the lane on its left that lands on the same pixel (a
// lane's (y0, x1) and (y1, x1) are its right neighbour's (y0, x0) and
// (y1, x0) wherever the flow is smooth), which halves them. A wave is one
// row of 64 columns: every lane stays active for the lane exchanges.
__global__ void __launch_bounds__(BX* BY)
    k_ad_photometric_bwd(const float* frames, const float2* flows, const float* scale, float* gframes,
                         float2* gflows, int T, int parity, int npairs, Dims d) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + wave_y(), z = blockIdx.z;
    if (y >= d.H) return;  // the whole wave
    const bool valid = x < d.W;
    const int xc = min(x, d.W - 1);
    const int b = z / npairs, t = 1 + parity + 2 * (z - b * npairs);
    const size_t cur = (size_t)(b * T + t) * d.H * d.W, prev = cur - (size_t)d.H * d.W, i = (size_t)y * d.W + xc;
    const Warp w = warp_at(xc, y, flows[cur + i], d);
    const Corners c = sample(frames + prev, w, d.W);
    const float sd = valid ? 2.0f * *scale * (frames[cur + i] - c.v) : 0.0f;
    float c00 = -sd * ((1.0f - w.wy) * (1.0f - w.wx)), c01 = -sd * ((1.0f - w.wy) * w.wx);
    float c10 = -sd * (w.wy * (1.0f - w.wx)), c11 = -sd * (w.wy * w.wx);
    // the lane on the left: its x1 column and rows, its right-hand terms
    const int v = valid ? 1 : 0;
    const int lv = lane_shift_int<true>(v), lx1 = lane_shift_int<true>(w.x1);
    const int ly0 = lane_shift_int<true>(w.y0), ly1 = lane_shift_int<true>(w.y1);
    const float l01 = lane_shift<true>(c01), l11 = lane_shift<true>(c11);
    // the lane on the right: whether it took this lane's right-hand terms
    const int rv = lane_shift_int<false>(v), rx0 = lane_shift_int<false>(w.x0);
    const int ry0 = lane_shift_int<false>(w.y0), ry1 = lane_shift_int<false>(w.y1);
    if (!valid) return;
    if (lv && lx1 == w.x0 && ly0 == w.y0) c00 += l01;
    if (lv && lx1 == w.x0 && ly1 == w.y1) c10 += l11;
    const bool gave0 = rv && rx0 == w.x1 && ry0 == w.y0, gave1 = rv && rx0 == w.x1 && ry1 == w.y1;
    gframes[cur + i] += sd;
    float* gp = gframes + prev;
    unsafeAtomicAdd(gp + (size_t)w.y0 * d.W + w.x0, c00);
    if (!gave0) unsafeAtomicAdd(gp + (size_t)w.y0 * d.W + w.x1, c01);
    unsafeAtomicAdd(gp + (size_t)w.y1 * d.W + w.x0, c10);
    if (!gave1) unsafeAtomicAdd(gp + (size_t)w.y1 * d.W + w.x1, c11);
    const float dvx = (1.0f - w.wy) * (c.p01 - c.p00) + w.wy * (c.p11 - c.p10);
    const float dvy = c.bot - c.top;
    float2 g = gflows[cur + i];
    if (w.fx) g.x += sd * dvx;
    if (w.fy) g.y += sd * dvy;
    gflows[cur + i] = g;
}

// Sum of n partials in double, in a fixed order (deterministic): thread t
// adds partials t, t + 256, t + 512, ... in sequence, then the 256 sums are
// combined pairwise (t += t + 128, then t += t + 64, ...). Thread 0 writes
// the total, or adds it to *out when acc is set.
constexpr int kSumThreads = 256;
__global__ void __launch_bounds__(kSumThreads) k_ad_sum_partials(const float* p, int n, double* out, int acc) {
    __shared__ double red[kSumThreads];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < n; i += kSumThreads) s += (double)p[i];
    red[t] = s;
    __syncthreads();
    for (int k = kSumThreads / 2; k > 0; k >>= 1) {
        if (t < k) red[t] += red[t + k];
        __syncthreads();
    }
    if (t == 0) *out = acc ? *out + red[0] : red[0];
}

}  // namespace

// Scratch of the allocation-free update adjoint (update_bwd_noalloc): the
// gamma partial sums of one call, one value per workgroup, 4 MiB. Each GPU
// holds its own copy of the array from the library's load until the
// process ends. In host code `scratch` names a host-side handle, not device
// memory: only the address scratch_ptr() looks up may be passed to kernels.
static __device__ __align__(256) float scratch[ad::kScratchPartials];

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
static float* scratch_ptr() {
    static std::mutex mu;
    static std::vector<float*> ptrs;
    int dev = 0;
    FF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void* p = nullptr;
        FF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw runtime_error("update adjoint: no device address for the scratch array");
        ptrs[dev] = (float*)p;
    }
    return ptrs[dev];
}

// ------------------------------------------------------------- launchers

namespace ad {

void image_model(const float* I, float* a0, float* g, int B, int H, int W, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_image_model, grid3(B, H, W), dim3(BX, BY), 0, s, I, a0, (float2*)g, Dims{H, W});
}
void image_model_bwd(const float* ga0, const float* gg, float* gI, float* work, int B, int H, int W, int acc,
                     hipStream_t s) {
    const size_t n = (size_t)B * H * W;
    hipLaunchKernelGGL(k_ad_image_model_bwd1, grid3(B, H, W), dim3(BX, BY), 0, s, ga0, (const float2*)gg, work,
                       work + n, Dims{H, W});
    hipLaunchKernelGGL(k_ad_image_model_bwd2, grid3(B, H, W), dim3(BX, BY), 0, s, work, work + n, gI, Dims{H, W},
                       acc);
}
void down(const float* I, float* O, float* work, int B, int H, int W, hipStream_t s) {
    const Dims d0{H, W}, d1{H, W / 2}, d2{H / 2, W / 2};
    hipLaunchKernelGGL((k_ad_down<0>), grid3(B, d1.H, d1.W), dim3(BX, BY), 0, s, I, work, d0, d1);
    hipLaunchKernelGGL((k_ad_down<1>), grid3(B, d2.H, d2.W), dim3(BX, BY), 0, s, work, O, d1, d2);
}
void down_bwd(const float* gO, float* gI, float* work, int B, int H, int W, int acc, hipStream_t s) {
    const Dims d0{H, W}, d1{H, W / 2}, d2{H / 2, W / 2};
    hipLaunchKernelGGL((k_ad_down_bwd<1>), grid3(B, d1.H, d1.W), dim3(BX, BY), 0, s, gO, work, d1, d2, 0);
    hipLaunchKernelGGL((k_ad_down_bwd<0>), grid3(B, d0.H, d0.W), dim3(BX, BY), 0, s, work, gI, d0, d1, acc);
}
void prop(int axis, const float* fin, const float* sin, const float* vin, float* fout, float* sout, float* vout, int B,
          int H, int W, float dt, int border, hipStream_t s) {
    const bool pay = sin != nullptr;
    const auto g = grid3(B, H, W);
    const dim3 blk(BX, BY);
    const Dims d{H, W};
    auto F = [](const float* p) { return (const float2*)p; };
    if (axis == 0) {
        if (pay) hipLaunchKernelGGL((k_ad_prop<0, true>), g, blk, 0, s, F(fin), sin, F(vin), (float2*)fout, sout, (float2*)vout, d, dt, border);
        else hipLaunchKernelGGL((k_ad_prop<0, false>), g, blk, 0, s, F(fin), sin, F(vin), (float2*)fout, sout, (float2*)vout, d, dt, border);
    } else {
        if (pay) hipLaunchKernelGGL((k_ad_prop<1, true>), g, blk, 0, s, F(fin), sin, F(vin), (float2*)fout, sout, (float2*)vout, d, dt, border);
        else hipLaunchKernelGGL((k_ad_prop<1, false>), g, blk, 0, s, F(fin), sin, F(vin), (float2*)fout, sout, (float2*)vout, d, dt, border);
    }
}
void prop_pair(const float* fin, const float* sin, const float* vin, float* fmid, float* smid, float* vmid,
               float* fout, float* sout, float* vout, int B, int H, int W, float dt, int border, hipStream_t s) {
    const bool pay = sin != nullptr, save = fmid != nullptr;
    const dim3 g((W + 63) / 64, (H + kPairRows - 1) / kPairRows, B), blk(64, kPairRows + 2);
    const Dims d{H, W};
    auto F = [](const float* p) { return (const float2*)p; };
    auto O = [](float* p) { return (float2*)p; };
    if (pay && save)
        hipLaunchKernelGGL((k_ad_prop_pair<true, true>), g, blk, 0, s, F(fin), sin, F(vin), O(fmid), smid, O(vmid), O(fout), sout, O(vout), d, dt, border);
    else if (pay)
        hipLaunchKernelGGL((k_ad_prop_pair<true, false>), g, blk, 0, s, F(fin), sin, F(vin), O(fmid), smid, O(vmid), O(fout), sout, O(vout), d, dt, border);
    else if (save)
        hipLaunchKernelGGL((k_ad_prop_pair<false, true>), g, blk, 0, s, F(fin), sin, F(vin), O(fmid), smid, O(vmid), O(fout), sout, O(vout), d, dt, border);
    else
        hipLaunchKernelGGL((k_ad_prop_pair<false, false>), g, blk, 0, s, F(fin), sin, F(vin), O(fmid), smid, O(vmid), O(fout), sout, O(vout), d, dt, border);
}
void prop_quad(const float* fin, const float* sin, const float* vin, float* fout, float* sout, float* vout, int B, int H,
               int W, float dt, int border, hipStream_t s) {
    constexpr int NW = kQuadWaves, RPW = kQuadRowsPerWave;
    const dim3 g((W + kQuadCols - 1) / kQuadCols, (H + kQuadRows - 1) / kQuadRows, B), blk(64, NW);
    const Dims d{H, W};
    if (sin)
        hipLaunchKernelGGL((k_ad_prop_quad<true, NW, RPW>), g, blk, 0, s, (const float2*)fin, sin, (const float2*)vin,
                           (float2*)fout, sout, (float2*)vout, d, dt, border);
    else
        hipLaunchKernelGGL((k_ad_prop_quad<false, NW, RPW>), g, blk, 0, s, (const float2*)fin, sin,
                           (const float2*)vin, (float2*)fout, sout, (float2*)vout, d, dt, border);
}
void prop_quad_bwd(const float* fin, const float* sin, const float* vin, const float* gfo, const float* gso,
                   const float* gvo, float* gfi, float* gsi, float* gvi, int B, int H, int W, float dt, int border,
                   int acc, hipStream_t s) {
    constexpr int NW = kQuadBwdWaves, RPW = kQuadBwdRowsPerWave;
    constexpr int rows = QuadBwdShape<NW, RPW>::rows;
    const dim3 g((W + kQuadBwdCols - 1) / kQuadBwdCols, (H + rows - 1) / rows, B), blk(64, NW);
    const Dims d{H, W};
    auto F = [](const float* p) { return (const float2*)p; };
    auto O = [](float* p) { return (float2*)p; };
    if (sin)
        hipLaunchKernelGGL((k_ad_prop_quad_bwd<true, NW, RPW>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso,
                           F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
    else
        hipLaunchKernelGGL((k_ad_prop_quad_bwd<false, NW, RPW>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso,
                           F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
}
void prop_pair_bwd(const float* fin, const float* sin, const float* vin, const float* fmid, const float* smid,
                   const float* vmid, const float* gfo, const float* gso, const float* gvo, float* gfi, float* gsi,
                   float* gvi, int B, int H, int W, float dt, int border, int acc, hipStream_t s) {
    const Dims d{H, W};
    auto F = [](const float* p) { return (const float2*)p; };
    auto O = [](float* p) { return (float2*)p; };
    if (!fmid) {  // recompute the X pass's result
        const dim3 g((W + kPairRcCols - 1) / kPairRcCols, (H + kPairBwdRows - 1) / kPairBwdRows, B),
            blk(64, kPairBwdRows);
        if (sin)
            hipLaunchKernelGGL((k_ad_prop_pair_bwd_rc<true>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
        else
            hipLaunchKernelGGL((k_ad_prop_pair_bwd_rc<false>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
        return;
    }
    const dim3 g((W + kPairCols - 1) / kPairCols, (H + kPairBwdRows - 1) / kPairBwdRows, B), blk(64, kPairBwdRows);
    if (sin)
        hipLaunchKernelGGL((k_ad_prop_pair_bwd<true>), g, blk, 0, s, F(fin), sin, F(vin), F(fmid), smid, F(vmid), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
    else
        hipLaunchKernelGGL((k_ad_prop_pair_bwd<false>), g, blk, 0, s, F(fin), sin, F(vin), F(fmid), smid, F(vmid), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
}
void prop_bwd(int axis, const float* fin, const float* sin, const float* vin, const float* gfo, const float* gso,
              const float* gvo, float* gfi, float* gsi, float* gvi, int B, int H, int W, float dt, int border,
              int acc, hipStream_t s) {
    const bool pay = sin != nullptr;
    const auto g = grid3(B, H, W);
    const dim3 blk(BX, BY);
    const Dims d{H, W};
    auto F = [](const float* p) { return (const float2*)p; };
    auto O = [](float* p) { return (float2*)p; };
    if (axis == 0) {
        if (pay) hipLaunchKernelGGL((k_ad_prop_bwd<0, true>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
        else hipLaunchKernelGGL((k_ad_prop_bwd<0, false>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
    } else {
        if (pay) hipLaunchKernelGGL((k_ad_prop_bwd<1, true>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
        else hipLaunchKernelGGL((k_ad_prop_bwd<1, false>), g, blk, 0, s, F(fin), sin, F(vin), F(gfo), gso, F(gvo), O(gfi), gsi, O(gvi), d, dt, border, acc);
    }
}
int blocks(int H, int W) { return ((W + BX - 1) / BX) * ((H + BY - 1) / BY); }
void update(const float* a0, const float* g, const float* aold, const float* f, float* fout, int B, int H, int W,
            float gamma, float maxflow, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_update, grid3(B, H, W), dim3(BX, BY), 0, s, a0, (const float2*)g, aold, (const float2*)f,
                       (float2*)fout, Dims{H, W}, gamma, maxflow);
}
void update_bwd(const float* a0, const float* g, const float* aold, const float* f, const float* gout, float* ga0,
                float* gg, float* gaold, float* gf, float* gamma_partial, int B, int H, int W, float gamma,
                float maxflow, int acc, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_update_bwd, grid3(B, H, W), dim3(BX, BY), 0, s, a0, (const float2*)g, aold,
                       (const float2*)f, (const float2*)gout, ga0, (float2*)gg, gaold, (float2*)gf, gamma_partial,
                       Dims{H, W}, gamma, maxflow, acc);
}
// The update adjoint with its gamma partial sums in `scratch` (allocates
// nothing), reduced into *g_gamma. The partials are all rewritten by the
// first launch, so nothing needs clearing between calls.
void update_bwd_noalloc(const float* a0, const float* g, const float* aold, const float* f, const float* gout,
                        float* ga0, float* gg, float* gaold, float* gf, double* g_gamma, int B, int H, int W,
                        float gamma, float maxflow, int acc, hipStream_t s) {
    const long long n = (long long)B * blocks(H, W);
    if (n > kScratchPartials)
        throw invalid_argument("ff_ad_update_backward_noalloc: B * ff_ad_partial_count(H, W) = " + std::to_string(n) +
                               " exceeds the scratch array's " + std::to_string(kScratchPartials) + " partial sums");
    float* part = scratch_ptr();
    update_bwd(a0, g, aold, f, gout, ga0, gg, gaold, gf, part, B, H, W, gamma, maxflow, acc & 15, s);
    hipLaunchKernelGGL(k_ad_sum_partials, dim3(1), dim3(kSumThreads), 0, s, (const float*)part, (int)n, g_gamma,
                       (acc >> 4) & 1);
}
void delta_update(const float* a0, const float* g, const float* sprop, const float* dprop, const float* upper,
                  const int* xmap, const int* ymap, float* dout, float* fout, int B, int H, int W, int Hu, int Wu,
                  float gamma, float maxflow, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_delta_update, grid3(B, H, W), dim3(BX, BY), 0, s, a0, (const float2*)g, sprop,
                       (const float2*)dprop, (const float2*)upper, xmap, ymap, (float2*)dout, (float2*)fout,
                       Dims{H, W}, Dims{Hu, Wu}, gamma, maxflow);
}
void delta_update_bwd(const float* a0, const float* g, const float* sprop, const float* dprop, const float* upper,
                      const int* xmap, const int* ymap, const float* gdout, const float* gfout, float* ga0, float* gg,
                      float* gsprop, float* gdprop, float* gup, float* gamma_partial, int B, int H, int W, int Hu,
                      int Wu, float gamma, float maxflow, int acc, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_delta_update_bwd, grid3(B, H, W), dim3(BX, BY), 0, s, a0, (const float2*)g, sprop,
                       (const float2*)dprop, (const float2*)upper, xmap, ymap, (const float2*)gdout,
                       (const float2*)gfout, ga0, (float2*)gg, gsprop, (float2*)gdprop, (float2*)gup, gamma_partial,
                       Dims{H, W}, Dims{Hu, Wu}, gamma, maxflow, acc);
}
void upper_gather(const float* gup, const int* xlo, const int* xhi, const int* ylo, const int* yhi, float* gupper,
                  int B, int H, int W, int Hu, int Wu, int acc, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_upper_gather, grid3(B, Hu, Wu), dim3(BX, BY), 0, s, (const float2*)gup, xlo, xhi, ylo, yhi,
                       (float2*)gupper, Dims{H, W}, Dims{Hu, Wu}, acc);
}
void smooth(int axis, const float* in, float* out, int B, int H, int W, hipStream_t s) {
    if (axis == 0) hipLaunchKernelGGL((k_ad_smooth<0>), grid3(B, H, W), dim3(BX, BY), 0, s, (const float2*)in, (float2*)out, Dims{H, W});
    else hipLaunchKernelGGL((k_ad_smooth<1>), grid3(B, H, W), dim3(BX, BY), 0, s, (const float2*)in, (float2*)out, Dims{H, W});
}
void smooth_pair(const float* in, float* out, int B, int H, int W, hipStream_t s) {
    const dim3 g((W + 63) / 64, (H + kSmoothRows - 1) / kSmoothRows, B), blk(64, kSmoothRows + 4);
    hipLaunchKernelGGL(k_ad_smooth_pair, g, blk, 0, s, (const float2*)in, (float2*)out, Dims{H, W});
}
void smooth_pair_bwd(const float* gout, const float* gout2, float* gin, int B, int H, int W, int acc, hipStream_t s) {
    const dim3 g((W + kSmoothBwdCols - 1) / kSmoothBwdCols, (H + kSmoothBwdRows - 1) / kSmoothBwdRows, B),
        blk(64, kSmoothBwdRows);
    hipLaunchKernelGGL(k_ad_smooth_pair_bwd, g, blk, 0, s, (const float2*)gout, (const float2*)gout2, (float2*)gin,
                       Dims{H, W}, acc);
}
void smooth_bwd(int axis, const float* gout, const float* gout2, float* gin, int B, int H, int W, int acc,
                hipStream_t s) {
    auto F = [](const float* p) { return (const float2*)p; };
    if (axis == 0) hipLaunchKernelGGL((k_ad_smooth_bwd<0>), grid3(B, H, W), dim3(BX, BY), 0, s, F(gout), F(gout2), (float2*)gin, Dims{H, W}, acc);
    else hipLaunchKernelGGL((k_ad_smooth_bwd<1>), grid3(B, H, W), dim3(BX, BY), 0, s, F(gout), F(gout2), (float2*)gin, Dims{H, W}, acc);
}

int photometric_partials(int B, int T, int H, int W) { return B * (T - 1) * blocks(H, W); }
void photometric(const float* frames, const float* flows, float* partial, int B, int T, int H, int W, hipStream_t s) {
    hipLaunchKernelGGL(k_ad_photometric, grid3(B * (T - 1), H, W), dim3(BX, BY), 0, s, frames, (const float2*)flows,
                       partial, T, Dims{H, W});
}
void photometric_bwd(const float* frames, const float* flows, const float* scale, float* gframes, float* gflows, int B,
                     int T, int H, int W, hipStream_t s) {
    // pairs t = 1, 3, ... then t = 2, 4, ... (see k_ad_photometric_bwd)
    for (int parity = 0; parity < 2; ++parity) {
        const int npairs = (T - parity) / 2;
        if (npairs < 1) continue;
        hipLaunchKernelGGL(k_ad_photometric_bwd, grid3(B * npairs, H, W), dim3(BX, BY), 0, s, frames,
                           (const float2*)flows, scale, gframes, (float2*)gflows, T, parity, npairs, Dims{H, W});
    }
}

}  // namespace ad

}  // namespace ffamd
