// ff_kernels.hpp -- host-side launchers of the HIP kernels (gfx950).
//
// Every launcher takes plain device pointers, sizes and pitches plus the
// stream to enqueue on; none of them allocates, synchronizes or throws on
// the device side. Shapes are validated by the stage classes before a
// launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace ffamd {

// Device view of one pitched 2-D image (element type implied by the call).
struct DevImg {
    void* data = nullptr;
    int h = 0;
    int w = 0;
    size_t pitch = 0;  // bytes between rows
};

// ------------------------------------------------------------ stage kernels
// One launch per reference kernel family; used by the stage classes
// (ImageModel, FlowPropagator, ...) and by the split compute methods.

// imagePrefilter_k + imageModel_k: a0 (float) and gradient (float2) of a u8
// (normalized read) or float image.
void launch_image_model(const DevImg& img, bool u8, const DevImg& a0, const DevImg& grad,
                        hipStream_t s);

// scalarProductF2_k: out = scalar * in on float2 images.
void launch_scale_f2(const DevImg& in, float scalar, const DevImg& out, hipStream_t s);

// One upwind sub-pass (axis 0 = x, 1 = y) of the flow and optional payloads
// (pass nullptr scalar/vector views for the plain propagator).
void launch_propagate_pass(int axis, const DevImg& flow_in, const DevImg& flow_out,
                           const DevImg* scalar_in, const DevImg* scalar_out, const DevImg* vec_in,
                           const DevImg* vec_out, float dt, int border, hipStream_t s);

// flowUpdate_k. image_out may alias a0old (same pixel read then written).
void launch_update(const DevImg& a0, const DevImg& grad, const DevImg& a0old, const DevImg& flow_old,
                   const DevImg& image_out, const DevImg& flow_out, float gamma, float maxflow,
                   hipStream_t s);

// deltaFlowUpdate_k. xmap/ymap hold the upper-level texel index of every
// column/row (the normalized-coordinate point-sampling rule evaluated on the
// host).
void launch_delta_update(const DevImg& a0, const DevImg& grad, const DevImg& a0old,
                         const DevImg& dflow_old, const DevImg& upper, const int* xmap,
                         const int* ymap, const DevImg& image_out, const DevImg& dflow_out,
                         const DevImg& flow_out, float gamma, float maxflow, hipStream_t s);

// One box-smoothing pass (axis 0 = x, 1 = y) on float2 images.
void launch_smooth_pass(int axis, const DevImg& in, const DevImg& out, hipStream_t s);

// One pyramid half-resolution pass (axis 0 = x, 1 = y), u8 or float.
void launch_down_pass(int axis, bool u8, const DevImg& in, const DevImg& out, hipStream_t s);

// Byte copy of a pitched 2-D region as a kernel (usable under graph capture,
// unlike a memset/memcpy node). row_bytes bytes per row.
void launch_copy2d(const DevImg& src, const DevImg& dst, size_t row_bytes, hipStream_t s);

// Copies of several pitched 2-D byte regions in one launch (16-byte
// accesses where every row of the region allows them).
struct CopyRegion {
    const void* src;
    void* dst;
    size_t src_pitch, dst_pitch, row_bytes;
    int rows;
};
void launch_copy_regions(const CopyRegion* regions, int n, hipStream_t s);

// flowToColor_k: RGBA colour (4 bytes per pixel) of every flow vector, from
// the colour wheel image (u8, 4 channels) read at the flow normalised by
// maxflow (ff_display.hip).
void launch_flow_to_color(const DevImg& flow, const DevImg& wheel, float maxflow, const DevImg& color, hipStream_t s);
// The colour wheel table (256 x 256 RGBA, rows top to bottom) and its size.
constexpr int kWheelHeight = 256, kWheelWidth = 256, kWheelDepth = 4;
const unsigned char* colorwheel_rgba(size_t* bytes);

// Perspective camera intrinsics (perspectiveCamera,
// include/flowfilter/gpu/camera.h:20-25), layout-identical to ff_camera.
struct Camera {
    float alphaX, alphaY, centerX, centerY;
};
// rotationalOpticalFlow_k: the flow of a camera rotating with angular
// velocity (wx, wy, wz), every pixel of `flow` (float2) (ff_rotation.hip).
void launch_rotational_flow(const Camera& cam, float wx, float wy, float wz, const DevImg& flow, hipStream_t s);
// LaxWendroffX_k / LaxWendroffY_k: one pass (axis 0: x, 1: y) over an image
// of `channels` (1 or 4) floats per pixel, advected by `flow` (float2).
void launch_lax_wendroff_pass(int axis, int channels, const DevImg& flow, const DevImg& in, const DevImg& out, float dt,
                              hipStream_t s);

// ------------------------------------------------------------ fused frame
// One kernel launch runs every level of a frame (and, for batched filters,
// of many frames): each level is a job whose output tiles are workgroups of
// the launch. Within a tile the whole per-level recurrence runs in
// registers: image model, propagation (with payloads on delta levels),
// update and smoothing.

// How a job reads its image.
enum ImgMode : int {
    IMG_F32 = 0,      // float image of the level's size
    IMG_U8 = 1,       // u8 image of the level's size (normalized read)
    IMG_U8_DOWN = 2,  // u8 image of twice the level's size, reduced by one
                      // pyramid step (imageDown{X,Y}_uint8_k) while loading
    IMG_U8_DOWN2 = 3, // u8 image of four times the level's size, reduced by
                      // two pyramid steps while loading
};

// Parameters of one level of one frame.
struct LevelArgs {
    DevImg img;              // input image (see img_mode)
    int img_mode = IMG_U8;
    // previous state (read)
    DevImg flow_s_in;        // smoothed flow, float2
    DevImg a0_in;            // updated image (old brightness), float
    DevImg dflow_in;         // delta flow, float2 (delta levels only)
    // new state (written)
    DevImg flow_s_out;
    DevImg a0_out;
    DevImg dflow_out;        // delta levels only
    DevImg flow_upd_out;     // unsmoothed updated flow (may be empty on a top level)
    // upper level flow (delta levels only) and its texel index maps
    DevImg upper;
    const int* xmap = nullptr;
    const int* ymap = nullptr;
    float gamma = 1.0f, maxflow = 1.0f, dt = 1.0f;
    int prop_iters = 1, smooth_iters = 1, border = 3;
    int delta = 0;           // 1: DeltaFlowFilter level, 0: top level (FlowFilter)
    // Index (in the same launch) of the job whose unsmoothed flow this job
    // reads as `upper`, or -1 when `upper` is not written in this launch.
    // The job must come earlier in the list.
    int wait_on = -1;
    // k >= 0: `img` is a filter's input frame, which a launch may instead
    // pass in place as its frame k (FusedFrame::launch); -1: always `img`.
    int frame_slot = -1;
};

// A frame a fused launch reads in place: device u8 image, row pitch in
// bytes (multiples of 4; pitch * rows below 2^31).
struct FrameRef {
    const void* ptr = nullptr;
    size_t pitch = 0;
};
// Frames one launch can take in place (a batch of more streams copies the
// rest into their input images).
constexpr int kMaxLaunchFrames = 128;

// Count of the waits of a FusedFrame's launches that gave up after their
// spin limit (the results of those frames are not valid). The kernel counts
// in a device word; read() copies it on the stream of the launches and
// waits for that stream only.
struct WaitWatch {
    WaitWatch() = default;
    WaitWatch(const WaitWatch&) = delete;
    WaitWatch& operator=(const WaitWatch&) = delete;
    ~WaitWatch();
    const unsigned* word = nullptr;  // the kernel's count (device)
    hipStream_t stream = nullptr;    // stream of the launches that count
    unsigned long long base = 0;     // counts of words since replaced
    unsigned long long total = 0;    // base + word, as last read
    bool pending = false;            // launches with waits since the last read
    unsigned* host = nullptr;        // pinned copy of the word
    unsigned long long read();
    // The word is replaced by a new zero word: keep its count in base.
    void retire(const unsigned* next);
};

// One observer's view of a WaitWatch: check() raises FF_ERR_RUNTIME when the
// count grew since this observer's last check.
struct WaitCheck {
    std::shared_ptr<WaitWatch> watch;
    unsigned long long seen = 0;
    void follow(const std::shared_ptr<WaitWatch>& w);
    void check(const char* who);
};

// Returns false when the level's parameters are outside what the fused
// kernel handles (the caller then runs the stage kernels).
bool fused_level_supported(const LevelArgs& a);

// Device tables and completion counters of fused frame launches. prepare()
// builds the job table of a list of levels (uploading it, after waiting for
// the work on stream s, when it matches none of the two most recent tables),
// choosing the tile height by timing the candidates the first time it sees
// a set of level shapes; it must not run under stream capture. launch()
// only enqueues the kernel.
class FusedFrame {
public:
    FusedFrame();
    ~FusedFrame();
    FusedFrame(const FusedFrame&) = delete;
    FusedFrame& operator=(const FusedFrame&) = delete;
    FusedFrame(FusedFrame&&) noexcept;
    FusedFrame& operator=(FusedFrame&&) noexcept;

    // The first call for a set of level shapes times every candidate tiling
    // on stream s (except what setWaves / setTopRows fixed).
    int prepare(const std::vector<LevelArgs>& jobs, hipStream_t s);
    // frames (optional, nframes <= kMaxLaunchFrames): the jobs whose
    // frame_slot is k read frames[k] in place instead of their img (when
    // its ptr is not null).
    void launch(int table, hipStream_t s, const FrameRef* frames = nullptr, int nframes = 0);
    // A graph holding launches of this frame was replayed on the stream:
    // the next direct launch clears the completion counters first.
    void replayed();
    // Waves per workgroup (a tile's buffer is 64 columns by waves * rows
    // per thread rows): 0 = autotune (default), or 16, 15, 8, 5, 4. waves()
    // reports the fixed or tuned value (0 before tuning).
    void setWaves(int nw);
    int waves() const;
    // Rows per thread of the tiles of top levels reduced on load (a
    // pyramid's top level): 0 = autotune (default), 4 or 8. topRows()
    // reports the fixed or tuned value (0 before tuning).
    void setTopRows(int rows);
    int topRows() const;
    // Rows per thread of the other levels' tiles: 0 (autotune with the
    // tile height, the default; 8 when both choices above are fixed), 4
    // (16 or 15 waves per workgroup only) or 8. rows() reads the fixed or
    // tuned value.
    void setRows(int rows);
    int rows() const;
    // Interior tiles (inside the propagation border, away from the image
    // edges) of a 4-row tiling in a tile body of their own: 0 = autotune
    // with the tiling (default), 1 = on (needs 4 rows per thread on the
    // levels below the top), 2 = off. interior() reads the fixed or tuned
    // choice (0 before tuning).
    void setInterior(int mode);
    int interior() const;
    // Waits that gave up after the spin limit (a scheduling fault; results
    // of such a frame are not valid). Waits for the stream of the launches.
    unsigned long long timeouts() const;
    const std::shared_ptr<WaitWatch>& watch() const;
    // Test hook: spins > 0 makes the waits of later launches wait for a
    // count no launch reaches and give up after `spins` polls (every
    // kernel still ends); 0 restores the normal waits.
    void setWaitTest(int spins);
    int waitTest() const;
    // Workgroups of a prepared table.
    int blocks(int table) const;
    // Identity of the device memory a launch of this table uses; it changes
    // whenever that memory is replaced, so recorded launches (graphs) keyed
    // by it never refer to freed memory.
    unsigned long long token(int table) const;
    // Whether some job of the table waits for another (see WaitingLaunch).
    bool waits(int table) const;

private:
    struct Impl;
    Impl* impl_;
};

// ------------------------------------------------- differentiable form
// Batched forward stages on float images and their adjoints
// (ff_adjoint.hip). B contiguous images of H x W; vector fields are float2.
namespace ad {
void image_model(const float* I, float* a0, float* g, int B, int H, int W, hipStream_t s);
void image_model_bwd(const float* ga0, const float* gg, float* gI, float* work, int B, int H, int W, int acc,
                     hipStream_t s);
void down(const float* I, float* O, float* work, int B, int H, int W, hipStream_t s);
void down_bwd(const float* gO, float* gI, float* work, int B, int H, int W, int acc, hipStream_t s);
void prop(int axis, const float* fin, const float* sin, const float* vin, float* fout, float* sout, float* vout, int B,
          int H, int W, float dt, int border, hipStream_t s);
void prop_pair(const float* fin, const float* sin, const float* vin, float* fmid, float* smid, float* vmid,
               float* fout, float* sout, float* vout, int B, int H, int W, float dt, int border, hipStream_t s);
void prop_quad(const float* fin, const float* sin, const float* vin, float* fout, float* sout, float* vout, int B, int H,
               int W, float dt, int border, hipStream_t s);
void prop_quad_bwd(const float* fin, const float* sin, const float* vin, const float* gfo, const float* gso,
                   const float* gvo, float* gfi, float* gsi, float* gvi, int B, int H, int W, float dt, int border,
                   int acc, hipStream_t s);
void prop_pair_bwd(const float* fin, const float* sin, const float* vin, const float* fmid, const float* smid,
                   const float* vmid, const float* gfo, const float* gso, const float* gvo, float* gfi, float* gsi,
                   float* gvi, int B, int H, int W, float dt, int border, int acc, hipStream_t s);
void prop_bwd(int axis, const float* fin, const float* sin, const float* vin, const float* gfo, const float* gso,
              const float* gvo, float* gfi, float* gsi, float* gvi, int B, int H, int W, float dt, int border,
              int acc, hipStream_t s);
int blocks(int H, int W);
void update(const float* a0, const float* g, const float* aold, const float* f, float* fout, int B, int H, int W,
            float gamma, float maxflow, hipStream_t s);
void update_bwd(const float* a0, const float* g, const float* aold, const float* f, const float* gout, float* ga0,
                float* gg, float* gaold, float* gf, float* gamma_partial, int B, int H, int W, float gamma,
                float maxflow, int acc, hipStream_t s);
// Largest B * blocks(H, W) update_bwd_noalloc accepts (its scratch array).
constexpr int kScratchPartials = 1 << 20;
void update_bwd_noalloc(const float* a0, const float* g, const float* aold, const float* f, const float* gout,
                        float* ga0, float* gg, float* gaold, float* gf, double* g_gamma, int B, int H, int W,
                        float gamma, float maxflow, int acc, hipStream_t s);
void delta_update(const float* a0, const float* g, const float* sprop, const float* dprop, const float* upper,
                  const int* xmap, const int* ymap, float* dout, float* fout, int B, int H, int W, int Hu, int Wu,
                  float gamma, float maxflow, hipStream_t s);
void delta_update_bwd(const float* a0, const float* g, const float* sprop, const float* dprop, const float* upper,
                      const int* xmap, const int* ymap, const float* gdout, const float* gfout, float* ga0, float* gg,
                      float* gsprop, float* gdprop, float* gup, float* gamma_partial, int B, int H, int W, int Hu,
                      int Wu, float gamma, float maxflow, int acc, hipStream_t s);
void upper_gather(const float* gup, const int* xlo, const int* xhi, const int* ylo, const int* yhi, float* gupper,
                  int B, int H, int W, int Hu, int Wu, int acc, hipStream_t s);
void smooth(int axis, const float* in, float* out, int B, int H, int W, hipStream_t s);
void smooth_pair(const float* in, float* out, int B, int H, int W, hipStream_t s);
void smooth_pair_bwd(const float* gout, const float* gout2, float* gin, int B, int H, int W, int acc,
                     hipStream_t s);
void smooth_bwd(int axis, const float* gout, const float* gout2, float* gin, int B, int H, int W, int acc,
                hipStream_t s);
int photometric_partials(int B, int T, int H, int W);
void photometric(const float* frames, const float* flows, float* partial, int B, int T, int H, int W, hipStream_t s);
void photometric_bwd(const float* frames, const float* flows, const float* scale, float* gframes, float* gflows, int B,
                     int T, int H, int W, hipStream_t s);
}  // namespace ad

}  // namespace ffamd
