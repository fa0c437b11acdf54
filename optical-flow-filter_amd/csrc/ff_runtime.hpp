// ff_runtime.hpp -- host runtime of the library: errors, the library
// stream, pitched device images with shared ownership, event timing.
//
// Counterparts in the reference: GPUImage (src/gpu/image.cu:24-214,
// include/flowfilter/gpu/image.h:47-123), Stage timing
// (src/gpu/pipeline.cu:21-89), assertError (src/gpu/error.cu:15-24).
// Texture objects have no counterpart here: kernels read pitched memory
// with clamped indices.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/flowfilter_amd.h"
#include "ff_kernels.hpp"

namespace ffamd {

// Library error carrying the C ABI status code. The C entry points catch it
// and turn it into a status plus a thread-local message.
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define FF_HIP(expr)                                                   \
    do {                                                               \
        hipError_t ff_e_ = (expr);                                     \
        if (ff_e_ != hipSuccess) ::ffamd::throw_hip(ff_e_, #expr, __FILE__, __LINE__); \
    } while (0)

// Checks for a failed kernel launch (invalid configuration and the like).
void check_launch(const char* what);

inline Error invalid_argument(const std::string& m) { return Error(FF_ERR_INVALID_ARGUMENT, m); }
inline Error logic_error(const std::string& m) { return Error(FF_ERR_LOGIC, m); }
inline Error runtime_error(const std::string& m) { return Error(FF_ERR_RUNTIME, m); }

// ------------------------------------------------------------- streams
//
// The reference runs everything on stream 0 with a host wait after every
// stage (src/gpu/pipeline.cu:63-82). Here every filter and batch owns a
// stream, and everything else (standalone stages, images created outside a
// filter) uses the library stream, which a caller may replace
// (ff_set_stream). A call on a filter binds the filter's stream to the
// calling thread for its duration (StreamScope); stream() is the bound
// stream, or the library stream when none is bound. Objects built inside a
// bound scope (a pyramid's levels) share the bound stream.

// A stream the library enqueues work on: owned by the box, a caller's
// stream, or (for a filter added to a batch) the stream of another box.
class StreamBox {
public:
    StreamBox() = default;
    StreamBox(const StreamBox&) = delete;
    StreamBox& operator=(const StreamBox&) = delete;
    ~StreamBox();
    // A box with a new stream of its own (non-blocking, default priority).
    static std::shared_ptr<StreamBox> create();
    // The library stream's box.
    static const std::shared_ptr<StreamBox>& library();
    hipStream_t get() const;
    // Enqueue on `s` (a caller's stream, not destroyed by the box); null
    // returns to the box's own stream. Waits for the work already enqueued
    // on the previous stream. For the library box, call it while no other
    // thread enqueues work through the library.
    void use(hipStream_t s);
    // Follow `to` (a batch's box); null ends it. Waits for the work already
    // enqueued on the previous stream.
    void follow(const std::shared_ptr<StreamBox>& to);
    // An event recorded only on this box's stream (reused).
    hipEvent_t event();

private:
    void drain();
    hipStream_t own_ = nullptr;
    std::atomic<hipStream_t> user_{nullptr};
    std::shared_ptr<StreamBox> alias_;
    hipEvent_t ev_ = nullptr;
};

// Binds a box to the calling thread while in scope (nests).
class StreamScope {
public:
    explicit StreamScope(const std::shared_ptr<StreamBox>& b);
    ~StreamScope();
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;

private:
    std::shared_ptr<StreamBox> prev_;
};

// The bound box (the library box when none is bound) and its stream.
const std::shared_ptr<StreamBox>& current_box();
// Whether a box is bound to the calling thread.
bool stream_bound();
hipStream_t stream();
void set_stream(hipStream_t s);
// Waits for all work the library enqueued on the device.
void synchronize();

// Device allocation shared by every handle of one image. The pointer is
// mutable so that a filter can exchange the buffers of two same-shape
// images (ping-pong state) while every outstanding handle follows. `home`
// is the box bound when it was allocated: the image's own copies and
// clears run on its stream.
struct Storage {
    void* ptr = nullptr;
    size_t bytes = 0;
    std::shared_ptr<StreamBox> home;
    Storage() = default;
    Storage(const Storage&) = delete;
    Storage& operator=(const Storage&) = delete;
    ~Storage();
};

// Row pitch used for every image: rows start on 256-byte boundaries.
size_t pitch_for(int width, int depth, int itemSize);

// Pitched 2-D device image with value semantics and shared ownership of
// the buffer (GPUImage, include/flowfilter/gpu/image.h:47-123).
class Image {
public:
    Image() = default;
    Image(int height, int width, int depth, int itemSize);

    int height() const { return h_; }
    int width() const { return w_; }
    int depth() const { return d_; }
    int itemSize() const { return item_; }
    size_t pitch() const { return pitch_; }
    size_t rowBytes() const { return (size_t)w_ * d_ * item_; }
    void* data() const { return st_ ? st_->ptr : nullptr; }
    bool allocated() const { return st_ && st_->ptr; }
    DevImg view() const { return DevImg{data(), h_, w_, pitch_}; }

    // Host transfers wait for completion (the reference issues synchronous
    // cudaMemcpy2D, src/gpu/image.cu:105,133). They run on the image's home
    // stream, after the work already enqueued there.
    void upload(const ff_image_t& host);
    // The same on stream s (the caller orders s after the image's readers).
    void upload(const ff_image_t& host, hipStream_t s);
    void download(ff_image_t& host) const;
    // Device-memory transfers and clears are enqueued on the home stream.
    void upload_device(const void* dev, size_t pitch);
    void download_device(void* dev, size_t pitch) const;
    void copyFrom(const Image& src);
    void clear();
    // Stream of the image's home box (the library stream when unallocated).
    hipStream_t homeStream() const;
    const std::shared_ptr<StreamBox>& home() const;

    bool sameShape(const Image& o) const {
        return h_ == o.h_ && w_ == o.w_ && d_ == o.d_ && item_ == o.item_;
    }
    bool matches(const ff_image_t& img) const {
        return h_ == img.height && w_ == img.width && d_ == img.depth && (size_t)item_ == img.itemSize;
    }
    // Exchange the device buffers behind this image and `o` (same shape);
    // every handle of either image sees the exchange.
    void swapBuffers(Image& o);
    std::string shapeString() const;
    static std::string shapeString(const ff_image_t& img);

private:
    void allocate();
    int h_ = 0, w_ = 0, d_ = 0, item_ = 0;
    size_t pitch_ = 0;
    std::shared_ptr<Storage> st_;
};

// Orders work across streams for images another object owns: on
// construction the current stream waits for the work enqueued so far on
// each image's home stream (when that is another stream); on destruction
// each such home stream waits for the work enqueued meanwhile on the
// current stream (so later writes there cannot overtake this work's reads).
class CrossStream {
public:
    CrossStream(std::initializer_list<const Image*> imgs);
    ~CrossStream();
    CrossStream(const CrossStream&) = delete;
    CrossStream& operator=(const CrossStream&) = delete;

private:
    hipStream_t cur_;
    std::shared_ptr<StreamBox> cur_box_;
    std::vector<std::shared_ptr<StreamBox>> others_;
};

// Orders fused frame launches in which some tiles wait for others of the
// same launch (a pyramid's delta level spins until its top level's tiles
// have signalled) one after another on the device, whatever stream they
// are on. Inside one launch the producers' blocks come first, and each XCD
// dispatches its blocks in order, so a launch alone always drains; but two
// such launches side by side can each fill an XCD with waiting tiles while
// the other's producers are queued behind them there, and both waits then
// give up at their spin limit. Construct one around every such launch (a
// graph launch included): under a process-wide lock held until
// destruction, it makes the stream wait for the end of the previous such
// launch when that was on another stream. Launches from one stream are
// ordered by the stream itself, so nothing is marked while they stay on
// one: the end of a launch on a library stream (StreamBox) is marked by an
// event recorded there only when another stream launches (a record between
// two frames of one filter costs about 5 us). On a caller's stream, which
// may be destroyed at any time, the first switch away waits on the host for
// the device, and from then on every launch there records its end. Nested
// guards on one thread, and guards while the stream is capturing, do
// nothing; if the previous launch's stream is being captured, the switch
// waits for the device. Launches without waits are not ordered.
class WaitingLaunch {
public:
    WaitingLaunch(hipStream_t s, bool waits);
    ~WaitingLaunch();
    WaitingLaunch(const WaitingLaunch&) = delete;
    WaitingLaunch& operator=(const WaitingLaunch&) = delete;

private:
    hipStream_t s_ = nullptr;
    int dev_ = 0;
    bool counted_ = false, locked_ = false;
};

// Host frames for a filter's own input image, double buffered: load()
// copies a host frame into a pinned host buffer (on a few threads), then
// enqueues its transfer into the spare device buffer on a copy stream of
// its own, after the last work that read that buffer, and exchanges the two
// buffers behind the input image's handles; the input's home stream waits
// for the transfer. It returns once the host memory has been read, without
// waiting for the device, so the transfer overlaps the previous frame's
// compute(). readBy(s) records that the work enqueued so far on s reads the
// current buffer (call it after each compute()).
class InputFrames {
public:
    InputFrames();
    InputFrames(InputFrames&&) noexcept;
    InputFrames& operator=(InputFrames&&) noexcept;
    ~InputFrames();
    void load(Image& input, const ff_image_t& host);
    void readBy(hipStream_t s);

private:
    struct State;
    std::unique_ptr<State> st_;
};

// Small device array (index maps and similar).
class DeviceArray {
public:
    void assign(const void* host, size_t bytes);
    void* data() const { return st_ ? st_->ptr : nullptr; }

private:
    std::shared_ptr<Storage> st_;
};

// Kernel profiling: while a profile is open, every launch whose kernel name
// contains the profile's filter string is bracketed by a pair of HIP events
// on its stream; profile_end waits for them and returns the summed
// duration and the launch count.
void profile_begin(const std::string& filter);
void profile_end(double* total_ms, long long* count);
bool profiling_active();

class ProfileScope {
public:
    ProfileScope(const char* kernel, hipStream_t s);
    ~ProfileScope();
    ProfileScope(const ProfileScope&) = delete;
    ProfileScope& operator=(const ProfileScope&) = delete;

private:
    int slot_ = -1;
    hipStream_t s_ = nullptr;
};

// Start/stop event pair of one stage. elapsedTime waits only when asked.
class Timer {
public:
    Timer() = default;
    Timer(const Timer&) = delete;
    Timer& operator=(const Timer&) = delete;
    Timer(Timer&& o) noexcept { *this = std::move(o); }
    Timer& operator=(Timer&& o) noexcept;
    ~Timer();
    void start();
    void stop();
    float elapsed();

private:
    void ensure();
    hipEvent_t start_ = nullptr, stop_ = nullptr;
    bool pending_ = false;
    float last_ = 0.0f;
};

}  // namespace ffamd
