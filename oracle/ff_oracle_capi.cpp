// ff_oracle_capi.cpp -- C entry points over ff_oracle.hpp for ctypes.
//
// TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library. Arrays are dense row-major (h, w[, c]); every function
// has an _f32 (parity oracle) and an _f64 (FP64 oracle) variant. Inputs of
// the FP64 variants are float32 arrays (the same data the GPU sees),
// outputs are float64.
#include "ff_oracle.hpp"

#include <cstring>

using namespace ffo;

namespace {

template <class T, class S>
Field<T> load(const S* p, int h, int w, int c) {
    Field<T> f(h, w, c);
    for (size_t i = 0; i < f.v.size(); ++i) f.v[i] = T(p[i]);
    return f;
}
template <class T, class D>
void store(const Field<T>& f, D* p) {
    for (size_t i = 0; i < f.v.size(); ++i) p[i] = D(f.v[i]);
}
Field<uint8_t> load_u8(const uint8_t* p, int h, int w) {
    Field<uint8_t> f(h, w, 1);
    std::memcpy(f.v.data(), p, f.v.size());
    return f;
}

}  // namespace

extern "C" {

void ffo_set_threads(int n) { num_threads() = n < 1 ? 1 : n; }
int ffo_get_threads() { return num_threads(); }

// Planted mistake (ffo::Mutation; 0 = the faithful oracle).
void ffo_set_mutation(int m) { mutation() = m; }
int ffo_get_mutation() { return mutation(); }

// ---- image model: u8 or f32 input --------------------------------------
#define IMAGE_MODEL(SUF, T)                                                                         \
    void ffo_image_model_u8_##SUF(int h, int w, const uint8_t* img, T* a0, T* grad) {              \
        Field<T> A, G;                                                                              \
        image_model<T>(load_u8(img, h, w), A, G);                                                   \
        store(A, a0);                                                                               \
        store(G, grad);                                                                             \
    }                                                                                               \
    void ffo_image_model_f32_##SUF(int h, int w, const float* img, T* a0, T* grad) {               \
        Field<T> A, G;                                                                              \
        image_model<T>(load<float>(img, h, w, 1), A, G);                                            \
        store(A, a0);                                                                               \
        store(G, grad);                                                                             \
    }
IMAGE_MODEL(f32, float)
IMAGE_MODEL(f64, double)

// ---- propagation ---------------------------------------------------------
#define PROPAGATE(SUF, T)                                                                           \
    void ffo_propagate_##SUF(int h, int w, const float* flow, int iterations, int border,          \
                             int invert, T* out) {                                                  \
        store(propagate<T>(load<T>(flow, h, w, 2), iterations, border, invert != 0), out);          \
    }                                                                                               \
    void ffo_propagate_pass_##SUF(int axis, int h, int w, const float* flow, const float* scalar,  \
                                  const float* vec, float dt, int border, T* flow_out,             \
                                  T* scalar_out, T* vec_out) {                                      \
        Field<T> F = load<T>(flow, h, w, 2), S, V, Fo, So, Vo;                                      \
        if (scalar) S = load<T>(scalar, h, w, 1);                                                   \
        if (vec) V = load<T>(vec, h, w, 2);                                                         \
        propagate_pass<T>(axis, F, Fo, scalar ? &S : nullptr, &So, vec ? &V : nullptr, &Vo, T(dt), \
                          border);                                                                  \
        store(Fo, flow_out);                                                                        \
        if (scalar) store(So, scalar_out);                                                          \
        if (vec) store(Vo, vec_out);                                                                \
    }                                                                                               \
    void ffo_propagate_payload_##SUF(int h, int w, const float* flow, const float* scalar,         \
                                     const float* vec, int iterations, int border, T* flow_out,    \
                                     T* scalar_out, T* vec_out) {                                   \
        Field<T> Fo, So, Vo;                                                                        \
        propagate_payload<T>(load<T>(flow, h, w, 2), load<T>(scalar, h, w, 1),                      \
                             load<T>(vec, h, w, 2), iterations, border, Fo, So, Vo);                \
        store(Fo, flow_out);                                                                        \
        store(So, scalar_out);                                                                      \
        store(Vo, vec_out);                                                                         \
    }
PROPAGATE(f32, float)
PROPAGATE(f64, double)

// ---- update -------------------------------------------------------------
#define UPDATE(SUF, T)                                                                              \
    void ffo_update_##SUF(int h, int w, const float* a0, const float* grad, const float* a0old,    \
                          const float* flow_old, float gamma, float maxflow, T* flow_out,           \
                          T* image_out) {                                                           \
        Field<T> Fo, Io;                                                                            \
        update<T>(load<T>(a0, h, w, 1), load<T>(grad, h, w, 2), load<T>(a0old, h, w, 1),            \
                  load<T>(flow_old, h, w, 2), T(gamma), T(maxflow), Fo, Io);                        \
        store(Fo, flow_out);                                                                        \
        store(Io, image_out);                                                                       \
    }                                                                                               \
    void ffo_delta_update_##SUF(int h, int w, const float* a0, const float* grad,                  \
                                const float* a0old, const float* dflow_old, int hu, int wu,        \
                                const float* upper, float gamma, float maxflow, T* dflow_out,      \
                                T* flow_out, T* image_out) {                                        \
        Field<T> Do, Fo, Io;                                                                        \
        delta_update<T>(load<T>(a0, h, w, 1), load<T>(grad, h, w, 2), load<T>(a0old, h, w, 1),     \
                        load<T>(dflow_old, h, w, 2), load<T>(upper, hu, wu, 2), T(gamma),           \
                        T(maxflow), Do, Fo, Io);                                                    \
        store(Do, dflow_out);                                                                       \
        store(Fo, flow_out);                                                                        \
        store(Io, image_out);                                                                       \
    }                                                                                               \
    int ffo_upper_index_##SUF(int x, int w, int wu) { return upper_index(x, w, wu, T(0)); }
UPDATE(f32, float)
UPDATE(f64, double)

// ---- smoothing ------------------------------------------------------------
#define SMOOTH(SUF, T)                                                                              \
    void ffo_smooth_##SUF(int h, int w, const float* flow, int iterations, T* out) {               \
        store(smooth<T>(load<T>(flow, h, w, 2), iterations), out);                                 \
    }
SMOOTH(f32, float)
SMOOTH(f64, double)

// ---- pyramid ----------------------------------------------------------------
// Writes levels 1..levels-1 back to back into `out` (level sizes by integer
// halving of the previous level).
#define PYRAMID(SUF, T)                                                                             \
    void ffo_pyramid_u8_##SUF(int h, int w, int levels, const uint8_t* img, uint8_t* out) {        \
        auto p = pyramid<T, uint8_t>(load_u8(img, h, w), levels);                                   \
        for (int l = 1; l < levels; ++l) {                                                          \
            std::memcpy(out, p[l].v.data(), p[l].v.size());                                         \
            out += p[l].v.size();                                                                   \
        }                                                                                           \
    }                                                                                               \
    void ffo_pyramid_f32_##SUF(int h, int w, int levels, const float* img, float* out) {           \
        auto p = pyramid<T, float>(load<float>(img, h, w, 1), levels);                              \
        for (int l = 1; l < levels; ++l) {                                                          \
            std::memcpy(out, p[l].v.data(), p[l].v.size() * sizeof(float));                         \
            out += p[l].v.size();                                                                   \
        }                                                                                           \
    }                                                                                               \
    uint8_t ffo_down_u8_##SUF(uint8_t km, uint8_t k0, uint8_t kp) { return down_u8(km, k0, kp, T(0)); }
PYRAMID(f32, float)
PYRAMID(f64, double)

// ---- display (FlowToColor) ---------------------------------------------------
void ffo_flow_to_color(int h, int w, const float* flow, float maxflow, int wh, int ww, const uint8_t* wheel,
                       uint8_t* out) {
    Field<uint8_t> wh4(wh, ww, 4), o;
    std::memcpy(wh4.v.data(), wheel, wh4.v.size());
    flow_to_color(load<float>(flow, h, w, 2), maxflow, wh4, o);
    std::memcpy(out, o.v.data(), o.v.size());
}
int ffo_wheel_texel_index(float c, int n) { return wheel_texel_index(c, n); }

// ---- camera, rotational flow, Lax-Wendroff ------------------------------------
void ffo_create_perspective_camera(float focal, int h, int w, float sensor_h, float sensor_w, float* out4) {
    const Camera c = create_perspective_camera(focal, h, w, sensor_h, sensor_w);
    out4[0] = c.alphaX;
    out4[1] = c.alphaY;
    out4[2] = c.centerX;
    out4[3] = c.centerY;
}
void ffo_rotational_flow(const float* cam4, float wx, float wy, float wz, int h, int w, float* out) {
    Field<float> f;
    rotational_flow(Camera{cam4[0], cam4[1], cam4[2], cam4[3]}, wx, wy, wz, h, w, f);
    store(f, out);
}
void ffo_lax_wendroff(int h, int w, int c, const float* flow, const float* img, int iterations, float* out) {
    store(lax_wendroff(load<float>(flow, h, w, 2), load<float>(img, h, w, c), iterations), out);
}
void ffo_lax_wendroff_pass(int axis, int h, int w, int c, const float* flow, const float* img, float dt, float* out) {
    Field<float> o;
    lax_wendroff_pass(axis, load<float>(flow, h, w, 2), load<float>(img, h, w, c), o, dt);
    store(o, out);
}

// ---- single-level filter (FlowFilter) ------------------------------------
#define FLOWFILTER(SUF, T)                                                                          \
    void* ffo_flowfilter_create_##SUF(int h, int w, int u8, int smooth_iters, float maxflow,       \
                                      float gamma, int configure_only) {                            \
        auto* f = new TopLevel<T>();                                                                \
        f->init(h, w, u8 != 0);                                                                     \
        if (!configure_only) {                                                                      \
            f->set_gamma(gamma);                                                                    \
            f->set_maxflow(maxflow);                                                                \
            f->smooth_iters = smooth_iters;                                                         \
        }                                                                                           \
        return f;                                                                                   \
    }                                                                                               \
    void ffo_flowfilter_destroy_##SUF(void* p) { delete (TopLevel<T>*)p; }                         \
    void ffo_flowfilter_set_gamma_##SUF(void* p, float g) { ((TopLevel<T>*)p)->set_gamma(g); }     \
    void ffo_flowfilter_set_maxflow_##SUF(void* p, float m) { ((TopLevel<T>*)p)->set_maxflow(m); } \
    void ffo_flowfilter_set_smooth_iterations_##SUF(void* p, int n) {                               \
        ((TopLevel<T>*)p)->smooth_iters = n;                                                        \
    }                                                                                               \
    void ffo_flowfilter_set_border_##SUF(void* p, int b) { ((TopLevel<T>*)p)->border = b; }        \
    void ffo_flowfilter_compute_u8_##SUF(void* p, const uint8_t* img) {                            \
        auto* f = (TopLevel<T>*)p;                                                                  \
        f->compute(load_u8(img, f->h, f->w));                                                       \
    }                                                                                               \
    void ffo_flowfilter_compute_f32_##SUF(void* p, const float* img) {                             \
        auto* f = (TopLevel<T>*)p;                                                                  \
        f->compute(load<float>(img, f->h, f->w, 1));                                                \
    }                                                                                               \
    void ffo_flowfilter_get_##SUF(void* p, int which, T* out) {                                    \
        auto* f = (TopLevel<T>*)p;                                                                  \
        const Field<T>* src[] = {&f->smoothed, &f->flow_upd, &f->image_upd, &f->a0, &f->grad,      \
                                 &f->prop_flow};                                                    \
        store(*src[which], out);                                                                    \
    }
FLOWFILTER(f32, float)
FLOWFILTER(f64, double)

// ---- standalone delta filter (DeltaFlowFilter::compute) ------------------
// The upper-level flow (hu x wu x 2) is passed with every compute, as the
// flow the reference filter's inputFlow image holds at that point.
#define DELTAFILTER(SUF, T)                                                                         \
    void* ffo_deltafilter_create_##SUF(int h, int w, int u8) {                                     \
        auto* f = new DeltaLevel<T>();                                                              \
        f->init(h, w, u8 != 0);                                                                     \
        return f;                                                                                   \
    }                                                                                               \
    void ffo_deltafilter_destroy_##SUF(void* p) { delete (DeltaLevel<T>*)p; }                      \
    void ffo_deltafilter_set_gamma_##SUF(void* p, float g) { ((DeltaLevel<T>*)p)->set_gamma(g); }  \
    void ffo_deltafilter_set_maxflow_##SUF(void* p, float m) { ((DeltaLevel<T>*)p)->set_maxflow(m); } \
    void ffo_deltafilter_set_smooth_iterations_##SUF(void* p, int n) {                              \
        ((DeltaLevel<T>*)p)->smooth_iters = n;                                                      \
    }                                                                                               \
    void ffo_deltafilter_set_border_##SUF(void* p, int b) { ((DeltaLevel<T>*)p)->border = b; }     \
    void ffo_deltafilter_compute_u8_##SUF(void* p, const uint8_t* img, int hu, int wu,             \
                                          const float* upper) {                                     \
        auto* f = (DeltaLevel<T>*)p;                                                                \
        f->compute(load_u8(img, f->h, f->w), load<T>(upper, hu, wu, 2));                            \
    }                                                                                               \
    void ffo_deltafilter_compute_f32_##SUF(void* p, const float* img, int hu, int wu,              \
                                           const float* upper) {                                    \
        auto* f = (DeltaLevel<T>*)p;                                                                \
        f->compute(load<float>(img, f->h, f->w, 1), load<T>(upper, hu, wu, 2));                     \
    }                                                                                               \
    /* which: 0 smoothed flow (getFlow), 1 updated flow, 2 updated image (getImage), */             \
    /* 3 delta flow, 4 propagated scalar payload                                     */             \
    void ffo_deltafilter_get_##SUF(void* p, int which, T* out) {                                   \
        auto* f = (DeltaLevel<T>*)p;                                                                \
        const Field<T>* src[] = {&f->smoothed, &f->flow_upd, &f->image_upd, &f->dflow_upd,         \
                                 &f->prop_scalar};                                                  \
        store(*src[which], out);                                                                    \
    }
DELTAFILTER(f32, float)
DELTAFILTER(f64, double)

// ---- pyramidal filter ----------------------------------------------------------
#define PYRAMIDAL(SUF, T)                                                                           \
    void* ffo_pyramidal_create_##SUF(int h, int w, int levels) {                                   \
        auto* f = new Pyramidal<T>();                                                               \
        f->init(h, w, levels);                                                                      \
        return f;                                                                                   \
    }                                                                                               \
    void ffo_pyramidal_destroy_##SUF(void* p) { delete (Pyramidal<T>*)p; }                         \
    void ffo_pyramidal_set_gamma_##SUF(void* p, int level, float g) {                               \
        ((Pyramidal<T>*)p)->set_gamma(level, g);                                                    \
    }                                                                                               \
    void ffo_pyramidal_set_maxflow_##SUF(void* p, float m) { ((Pyramidal<T>*)p)->set_maxflow(m); } \
    void ffo_pyramidal_set_smooth_iterations_##SUF(void* p, int level, int n) {                     \
        ((Pyramidal<T>*)p)->set_smooth_iterations(level, n);                                        \
    }                                                                                               \
    void ffo_pyramidal_set_border_##SUF(void* p, int b) { ((Pyramidal<T>*)p)->set_border(b); }     \
    void ffo_pyramidal_compute_##SUF(void* p, const uint8_t* img) {                                \
        auto* f = (Pyramidal<T>*)p;                                                                 \
        f->compute(load_u8(img, f->h, f->w));                                                       \
    }                                                                                               \
    /* which: 0 downloadFlow, 1 downloadImage, 2 getFlow */                                         \
    void ffo_pyramidal_get_##SUF(void* p, int which, T* out) {                                     \
        auto* f = (Pyramidal<T>*)p;                                                                 \
        const Field<T>& src = which == 0 ? f->flow() : (which == 1 ? f->image() : f->device_flow()); \
        store(src, out);                                                                            \
    }                                                                                               \
    /* per-level state: 0 smoothed flow, 1 updated (unsmoothed) flow, 2 updated image,          */  \
    /* 3 delta flow (non-top levels only)                                                        */  \
    void ffo_pyramidal_get_level_##SUF(void* p, int level, int which, T* out) {                    \
        auto* f = (Pyramidal<T>*)p;                                                                 \
        if (level == f->L - 1) {                                                                    \
            const Field<T>* src[] = {&f->top.smoothed, &f->top.flow_upd, &f->top.image_upd,        \
                                     &f->top.flow_upd};                                             \
            store(*src[which], out);                                                                \
        } else {                                                                                    \
            auto& d = f->low[level];                                                                \
            const Field<T>* src[] = {&d.smoothed, &d.flow_upd, &d.image_upd, &d.dflow_upd};        \
            store(*src[which], out);                                                                \
        }                                                                                           \
    }
PYRAMIDAL(f32, float)
PYRAMIDAL(f64, double)

}  // extern "C"

// ---- exhaustive checks of arithmetic identities the HIP kernels rely on ----
// (evaluated on the CPU; both count mismatches, 0 means the identity holds)
extern "C" {

// u8 normalized read: product with the rounded reciprocal plus one fused
// correction step vs the IEEE quotient v / 255.0f, for all 256 v.
int ffo_count_unorm8_fma_mismatch(void) {
    const float r = 1.0f / 255.0f;
    int bad = 0;
    for (int v = 0; v < 256; ++v) {
        const float f = (float)v;
        const float q = f * r;
        const float e = std::fmaf(-q, 255.0f, f);
        const float g = std::fmaf(e, r, q);
        if (g != unorm8<float>((uint8_t)v)) ++bad;
    }
    return bad;
}

// u8 pyramid tap: the reference's float/double expression vs the integer
// form (2*k0 + km + kp) >> 2, for all 2^24 (km, k0, kp).
long long ffo_count_down_u8_int_mismatch(void) {
    long long bad = 0;
    for (int km = 0; km < 256; ++km)
        for (int k0 = 0; k0 < 256; ++k0)
            for (int kp = 0; kp < 256; ++kp) {
                const uint8_t a = down_u8((uint8_t)km, (uint8_t)k0, (uint8_t)kp, 0.0f);
                const uint8_t b = (uint8_t)((2 * k0 + km + kp) >> 2);
                if (a != b) ++bad;
            }
    return bad;
}

}  // extern "C"
