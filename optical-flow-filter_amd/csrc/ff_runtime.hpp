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

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

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

// Library-wide stream (the reference runs everything on stream 0 of
// device 0). Created on first use on the current device; a caller may
// substitute its own stream.
hipStream_t stream();
void set_stream(hipStream_t s);
void synchronize();

// Device allocation shared by every handle of one image. The pointer is
// mutable so that a filter can exchange the buffers of two same-shape
// images (ping-pong state) while every outstanding handle follows.
struct Storage {
    void* ptr = nullptr;
    size_t bytes = 0;
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
    // cudaMemcpy2D, src/gpu/image.cu:105,133).
    void upload(const ff_image_t& host);
    void download(ff_image_t& host) const;
    // Device-memory transfers are enqueued on the library stream.
    void upload_device(const void* dev, size_t pitch);
    void download_device(void* dev, size_t pitch) const;
    void copyFrom(const Image& src);
    void clear();

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
