// ff_runtime.cpp -- see ff_runtime.hpp.
#include "ff_runtime.hpp"

#include "ff_kernels.hpp"

#include <sched.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <set>
#include <sstream>
#include <thread>
#include <vector>

namespace ffamd {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
    std::ostringstream os;
    os << "HIP error '" << hipGetErrorString(e) << "' (" << (int)e << ") from " << what << " at " << file << ":"
       << line;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) throw Error(FF_ERR_BAD_ALLOC, os.str());
    throw Error(FF_ERR_HIP, os.str());
}

void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw_hip(e, what, __FILE__, __LINE__);
}

// ------------------------------------------------------------------ stream

namespace {
// The streams of live StreamBoxes, which WaitingLaunch may mark after the
// fact (defined with WaitingLaunch below).
void owned_insert(hipStream_t s);
void owned_erase(hipStream_t s);
}  // namespace

StreamBox::~StreamBox() {
    if (ev_) (void)hipEventDestroy(ev_);
    if (own_) {
        // finish first: then no launch on it is left to order, and a new
        // stream may be given the same handle
        (void)hipStreamSynchronize(own_);
        owned_erase(own_);
        (void)hipStreamDestroy(own_);
    }
}

std::shared_ptr<StreamBox> StreamBox::create() {
    auto b = std::make_shared<StreamBox>();
    FF_HIP(hipStreamCreateWithFlags(&b->own_, hipStreamNonBlocking));
    owned_insert(b->own_);
    return b;
}

const std::shared_ptr<StreamBox>& StreamBox::library() {
    // kept for the process (never destroyed: static destructors may run
    // after the HIP runtime has shut down)
    static const std::shared_ptr<StreamBox>* lib = new std::shared_ptr<StreamBox>(create());
    return *lib;
}

hipStream_t StreamBox::get() const {
    if (alias_) return alias_->get();
    const hipStream_t u = user_.load(std::memory_order_acquire);
    return u ? u : own_;
}

void StreamBox::drain() { FF_HIP(hipStreamSynchronize(get())); }

void StreamBox::use(hipStream_t s) {
    drain();
    alias_.reset();
    user_.store(s, std::memory_order_release);
}

void StreamBox::follow(const std::shared_ptr<StreamBox>& to) {
    if (to.get() == this) return;
    drain();
    alias_ = to;
}

hipEvent_t StreamBox::event() {
    if (!ev_) FF_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
    return ev_;
}

namespace {
thread_local std::shared_ptr<StreamBox> t_bound;
}  // namespace

StreamScope::StreamScope(const std::shared_ptr<StreamBox>& b) : prev_(t_bound) {
    if (b) t_bound = b;
}
StreamScope::~StreamScope() { t_bound = prev_; }

const std::shared_ptr<StreamBox>& current_box() { return t_bound ? t_bound : StreamBox::library(); }

bool stream_bound() { return t_bound != nullptr; }

hipStream_t stream() { return current_box()->get(); }

void set_stream(hipStream_t s) { StreamBox::library()->use(s); }

void synchronize() { FF_HIP(hipDeviceSynchronize()); }

CrossStream::CrossStream(std::initializer_list<const Image*> imgs) : cur_(stream()), cur_box_(current_box()) {
    for (const Image* im : imgs) {
        if (!im || !im->allocated()) continue;
        const std::shared_ptr<StreamBox>& h = im->home();
        if (!h || h->get() == cur_) continue;
        bool seen = false;
        for (const auto& o : others_) seen = seen || o->get() == h->get();
        if (seen) continue;
        FF_HIP(hipEventRecord(h->event(), h->get()));
        FF_HIP(hipStreamWaitEvent(cur_, h->event(), 0));
        others_.push_back(h);
    }
}

CrossStream::~CrossStream() {
    if (others_.empty()) return;
    hipEvent_t e = cur_box_->event();
    if (hipEventRecord(e, cur_) != hipSuccess) return;
    for (const auto& o : others_) (void)hipStreamWaitEvent(o->get(), e, 0);
}

// ----------------------------------------------------------------- staging

namespace {
// Copies host images into pinned memory on a few threads. One thread reads
// pageable memory at ~20 GB/s (2048x1088 u8 frame: 0.1 ms), which is more
// than a frame's compute; up to eight split it (parts of at least 256 KiB).
// The helper threads are made on first use and live until the process
// ends (the object is never destroyed); a forked child copies on its own
// thread.
class HostCopier {
public:
    static HostCopier& get() {
        static HostCopier* c = new HostCopier();
        return *c;
    }

    // rows [0, h) of `row` bytes from src (pitch spitch) to dst (pitch dpitch)
    void copy(char* dst, size_t dpitch, const char* src, size_t spitch, size_t row, int h) {
        if (h <= 0 || row == 0) return;
        const bool linear = dpitch == spitch;
        const size_t span = linear ? (size_t)(h - 1) * dpitch + row : (size_t)h * row;
        const int parts = (int)std::min<size_t>((size_t)helpers_ + 1, std::max<size_t>(1, span / kMinPart));
        if (parts <= 1 || getpid() != pid_) {
            run(Job{dst, dpitch, src, spitch, row, h, linear, span, 1}, 0);
            return;
        }
        std::lock_guard<std::mutex> call(call_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = Job{dst, dpitch, src, spitch, row, h, linear, span, parts};
            pending_ = parts - 1;
            ++gen_;
        }
        wake_.notify_all();
        run(job_, 0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    static constexpr size_t kMinPart = 256 << 10;
    static constexpr int kMaxHelpers = 7;
    struct Job {
        char* dst;
        size_t dpitch;
        const char* src;
        size_t spitch, row;
        int h;
        bool linear;
        size_t span;
        int parts;
    };

    HostCopier() : pid_(getpid()) {
        cpu_set_t set;
        int cpus = 1;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
        helpers_ = std::max(0, std::min(kMaxHelpers, cpus / 2 - 1));
        for (int i = 0; i < helpers_; ++i) std::thread([this, i] { loop(i + 1); }).detach();
    }

    // part k of the job: a 4 KiB-aligned byte range of a linear copy, or a
    // range of rows
    static void run(const Job& j, int k) {
        if (j.linear) {
            const size_t per = ((j.span + j.parts - 1) / j.parts + 4095) & ~(size_t)4095;
            const size_t b = std::min(j.span, per * (size_t)k), e = std::min(j.span, b + per);
            if (e > b) std::memcpy(j.dst + b, j.src + b, e - b);
            return;
        }
        const int per = (j.h + j.parts - 1) / j.parts;
        const int r1 = std::min(j.h, per * (k + 1));
        for (int r = per * k; r < r1; ++r)
            std::memcpy(j.dst + (size_t)r * j.dpitch, j.src + (size_t)r * j.spitch, j.row);
    }

    void loop(int k) {
        unsigned long long seen = 0;
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                wake_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                j = job_;
            }
            if (k < j.parts) {
                run(j, k);
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }

    const pid_t pid_;
    int helpers_ = 0;
    std::mutex call_mu_, mu_;
    std::condition_variable wake_, done_;
    Job job_{};
    int pending_ = 0;
    unsigned long long gen_ = 0;
};

// Host <-> device copies of a whole image through pinned memory, in chunks
// of rows: the DMA of one chunk runs while the host threads copy the next
// (or previous) one between the pinned buffer and the host image, in either
// row layout. Two chunk buffers of about kChunk bytes and their events,
// made on first use and kept for the process; g_stage_mu serialises their
// users.
constexpr size_t kChunk = 4 << 20;
std::mutex g_stage_mu;
struct ChunkStage {
    char* buf[2] = {nullptr, nullptr};
    size_t bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
ChunkStage g_chunks;

// Rows per chunk for a device pitch, and the buffers sized for them (call
// with g_stage_mu held).
int chunk_rows(size_t pitch, size_t row_bytes) {
    const int rows = (int)std::max<size_t>(1, kChunk / pitch);
    const size_t need = (size_t)(rows - 1) * pitch + row_bytes;
    ChunkStage& c = g_chunks;
    if (c.bytes < need) {
        for (int k = 0; k < 2; ++k) {
            if (c.buf[k]) FF_HIP(hipHostFree(c.buf[k]));
            c.buf[k] = nullptr;
        }
        c.bytes = 0;
        for (int k = 0; k < 2; ++k) FF_HIP(hipHostMalloc((void**)&c.buf[k], need, hipHostMallocDefault));
        c.bytes = need;
    }
    for (int k = 0; k < 2; ++k)
        if (!c.ev[k]) FF_HIP(hipEventCreateWithFlags(&c.ev[k], hipEventDisableTiming));
    return rows;
}
}  // namespace

// ----------------------------------------------------------------- storage

Storage::~Storage() {
    if (ptr) (void)hipFree(ptr);
}

size_t pitch_for(int width, int depth, int itemSize) {
    const size_t row = (size_t)width * depth * itemSize;
    return (row + 255) / 256 * 256;
}

Image::Image(int height, int width, int depth, int itemSize) : h_(height), w_(width), d_(depth), item_(itemSize) {
    allocate();
}

void Image::allocate() {
    if (h_ < 0 || w_ < 0 || d_ < 0 || item_ < 0) throw invalid_argument("GPUImage: negative dimension " + shapeString());
    pitch_ = pitch_for(w_, d_, item_);
    st_ = std::make_shared<Storage>();
    st_->home = current_box();
    const size_t bytes = pitch_ * (size_t)h_;
    if (bytes > 0) {
        FF_HIP(hipMalloc(&st_->ptr, bytes));
        st_->bytes = bytes;
    }
}

const std::shared_ptr<StreamBox>& Image::home() const {
    return st_ && st_->home ? st_->home : StreamBox::library();
}

hipStream_t Image::homeStream() const { return home()->get(); }

std::string Image::shapeString() const {
    std::ostringstream os;
    os << "[" << h_ << ", " << w_ << ", " << d_ << "][" << item_ << "]";
    return os.str();
}

std::string Image::shapeString(const ff_image_t& img) {
    std::ostringstream os;
    os << "[" << img.height << ", " << img.width << ", " << img.depth << "][" << img.itemSize << "]";
    return os.str();
}

// GPUImage::upload (src/gpu/image.cu:81-121): an unallocated image takes the
// host image's shape; otherwise shapes must match.
void Image::upload(const ff_image_t& host) {
    if (!st_) {
        h_ = host.height;
        w_ = host.width;
        d_ = host.depth;
        item_ = (int)host.itemSize;
        allocate();
    }
    upload(host, homeStream());
}

void Image::upload(const ff_image_t& host, hipStream_t s) {
    if (!st_) {
        h_ = host.height;
        w_ = host.width;
        d_ = host.depth;
        item_ = (int)host.itemSize;
        allocate();
    }
    if (!matches(host))
        throw invalid_argument("GPUImage::upload(): shapes do not match. Required: " + shapeString() +
                               ", passed: " + shapeString(host));
    if (rowBytes() == 0 || h_ == 0) return;
    if (!host.data) throw invalid_argument("GPUImage::upload(): null host data");
    if (host.pitch < rowBytes()) throw invalid_argument("GPUImage::upload(): host pitch smaller than a row");
    // host rows -> pinned chunk (host threads) -> device (DMA on s), the
    // copy of chunk c+1 overlapping the DMA of chunk c
    std::lock_guard<std::mutex> lk(g_stage_mu);
    const int rows = chunk_rows(pitch_, rowBytes());
    ChunkStage& cs = g_chunks;
    for (int r0 = 0, c = 0; r0 < h_; r0 += rows, ++c) {
        const int k = c & 1, n = std::min(rows, h_ - r0);
        if (c >= 2) FF_HIP(hipEventSynchronize(cs.ev[k]));
        HostCopier::get().copy(cs.buf[k], pitch_, (const char*)host.data + (size_t)r0 * host.pitch, host.pitch,
                               rowBytes(), n);
        FF_HIP(hipMemcpyAsync((char*)data() + (size_t)r0 * pitch_, cs.buf[k], (size_t)(n - 1) * pitch_ + rowBytes(),
                              hipMemcpyHostToDevice, s));
        FF_HIP(hipEventRecord(cs.ev[k], s));
    }
    FF_HIP(hipStreamSynchronize(s));
}

// GPUImage::download (src/gpu/image.cu:123-150). The reference prints and
// returns on an unallocated image; here it is an error.
void Image::download(ff_image_t& host) const {
    if (!st_) throw logic_error("GPUImage::download(): unallocated image");
    if (!matches(host))
        throw invalid_argument("GPUImage::download(): shapes do not match. Required: " + shapeString() +
                               ", passed: " + shapeString(host));
    if (rowBytes() == 0 || h_ == 0) return;
    if (!host.data) throw invalid_argument("GPUImage::download(): null host data");
    if (host.pitch < rowBytes()) throw invalid_argument("GPUImage::download(): host pitch smaller than a row");
    // device -> pinned chunk (DMA on the home stream) -> host rows (host
    // threads), the DMA of chunk c+2 issued as soon as chunk c is copied out
    hipStream_t s = homeStream();
    std::lock_guard<std::mutex> lk(g_stage_mu);
    const int rows = chunk_rows(pitch_, rowBytes());
    ChunkStage& cs = g_chunks;
    const int n_chunks = (h_ + rows - 1) / rows;
    auto issue = [&](int c) {
        const int r0 = c * rows, n = std::min(rows, h_ - r0);
        FF_HIP(hipMemcpyAsync(cs.buf[c & 1], (const char*)data() + (size_t)r0 * pitch_,
                              (size_t)(n - 1) * pitch_ + rowBytes(), hipMemcpyDeviceToHost, s));
        FF_HIP(hipEventRecord(cs.ev[c & 1], s));
    };
    for (int c = 0; c < std::min(2, n_chunks); ++c) issue(c);
    for (int c = 0; c < n_chunks; ++c) {
        const int r0 = c * rows, n = std::min(rows, h_ - r0);
        FF_HIP(hipEventSynchronize(cs.ev[c & 1]));
        HostCopier::get().copy((char*)host.data + (size_t)r0 * host.pitch, host.pitch, cs.buf[c & 1], pitch_,
                               rowBytes(), n);
        if (c + 2 < n_chunks) issue(c + 2);
    }
}

void Image::upload_device(const void* dev, size_t pitch) {
    if (!st_) throw logic_error("GPUImage::upload_device(): unallocated image");
    if (rowBytes() == 0 || h_ == 0) return;
    if (!dev) throw invalid_argument("GPUImage::upload_device(): null device pointer");
    if (pitch < rowBytes()) throw invalid_argument("GPUImage::upload_device(): pitch smaller than a row");
    const CopyRegion r{dev, data(), pitch, pitch_, rowBytes(), h_};
    launch_copy_regions(&r, 1, homeStream());
    check_launch("GPUImage::upload_device");
}

void Image::download_device(void* dev, size_t pitch) const {
    if (!st_) throw logic_error("GPUImage::download_device(): unallocated image");
    if (rowBytes() == 0 || h_ == 0) return;
    if (!dev) throw invalid_argument("GPUImage::download_device(): null device pointer");
    if (pitch < rowBytes()) throw invalid_argument("GPUImage::download_device(): pitch smaller than a row");
    FF_HIP(hipMemcpy2DAsync(dev, pitch, data(), pitch_, rowBytes(), h_, hipMemcpyDeviceToDevice, homeStream()));
}

// GPUImage::copyFrom (src/gpu/image.cu:152-170): same shape and pitch
// required (compareShapeGPU, :207-214). Enqueued as a copy kernel so that it
// orders with the filter kernels on the library stream.
void Image::copyFrom(const Image& src) {
    if (!sameShape(src) || pitch_ != src.pitch_)
        throw invalid_argument("GPUImage::copyFrom(): shapes do not match. Required: " + shapeString() +
                               ", passed: " + src.shapeString());
    if (!st_ || !src.st_) throw logic_error("GPUImage::copyFrom(): unallocated image");
    if (rowBytes() == 0 || h_ == 0) return;
    StreamScope sc(home());
    CrossStream x{&src};
    launch_copy2d(src.view(), view(), rowBytes(), stream());
    check_launch("copyFrom");
}

// GPUImage::clear (src/gpu/image.cu:172-176).
void Image::clear() {
    if (!st_ || !st_->ptr) return;
    FF_HIP(hipMemset2DAsync(data(), pitch_, 0, rowBytes(), h_, homeStream()));
}

void Image::swapBuffers(Image& o) {
    if (!sameShape(o) || pitch_ != o.pitch_) throw logic_error("swapBuffers: shapes differ");
    if (!st_ || !o.st_) throw logic_error("swapBuffers: unallocated image");
    std::swap(st_->ptr, o.st_->ptr);
    std::swap(st_->bytes, o.st_->bytes);
}

// ------------------------------------------------------------ WaitingLaunch

namespace {
// What marks the end of a device's last waiting launch:
//  NONE:     nothing needs marking (no such launch, or its stream finished);
//  PENDING:  it ran on a live library stream and is not marked yet: an event
//            recorded on that stream when another stream launches marks it
//            (with whatever was queued there since, which only orders more);
//  RECORDED: `done` was recorded right after it;
//  UNMARKED: it ran on a caller's stream, which may since have been
//            destroyed, and was not marked.
enum class LastMark { NONE, PENDING, RECORDED, UNMARKED };
struct WaitingState {
    hipStream_t last = nullptr;  // the null stream is a stream too
    hipEvent_t done = nullptr;
    LastMark mark = LastMark::NONE;
    bool eager = false;  // record after every launch on a caller's stream
};
// Kept for the process (never destroyed), like the library stream: boxes
// may be destroyed during the process's exit.
struct Waiting {
    std::mutex mu;
    std::vector<WaitingState> dev;
    std::set<hipStream_t> owned;
};
Waiting& waiting() {
    static Waiting* w = new Waiting;
    return *w;
}
thread_local int t_waiting_depth = 0;

void owned_insert(hipStream_t s) {
    Waiting& g = waiting();
    std::lock_guard<std::mutex> lk(g.mu);
    g.owned.insert(s);
}
void owned_erase(hipStream_t s) {
    Waiting& g = waiting();
    std::lock_guard<std::mutex> lk(g.mu);
    g.owned.erase(s);
    for (WaitingState& w : g.dev)
        if (w.last == s && w.mark != LastMark::RECORDED) w.mark = LastMark::NONE;
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FF_HIP(hipStreamIsCapturing(s, &cs));
    return cs != hipStreamCaptureStatusNone;
}
}  // namespace

WaitingLaunch::WaitingLaunch(hipStream_t s, bool waits) : s_(s) {
    if (!waits) return;
    // counted (and so undone by the destructor) only once nothing can throw
    if (t_waiting_depth > 0 || capturing(s)) {
        ++t_waiting_depth;
        counted_ = true;
        return;
    }
    FF_HIP(hipGetDevice(&dev_));
    Waiting& g = waiting();
    std::unique_lock<std::mutex> lk(g.mu);
    if ((int)g.dev.size() <= dev_) g.dev.resize(dev_ + 1);
    WaitingState& w = g.dev[dev_];
    if (w.last != s && w.mark != LastMark::NONE) {
        if (!w.done) FF_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
        if (w.mark == LastMark::PENDING && !capturing(w.last)) {
            FF_HIP(hipEventRecord(w.done, w.last));
            FF_HIP(hipStreamWaitEvent(s, w.done, 0));
        } else if (w.mark == LastMark::RECORDED) {
            FF_HIP(hipStreamWaitEvent(s, w.done, 0));
        } else {
            // unmarked (or its stream is being captured): wait for the
            // device, and from now on mark launches on callers' streams as
            // they end
            FF_HIP(hipDeviceSynchronize());
            w.eager = true;
        }
    }
    lk.release();  // held until the destructor
    ++t_waiting_depth;
    counted_ = true;
    locked_ = true;
}

WaitingLaunch::~WaitingLaunch() {
    if (!counted_) return;
    --t_waiting_depth;
    if (!locked_) return;
    Waiting& g = waiting();
    WaitingState& w = g.dev[dev_];
    w.last = s_;
    if (g.owned.count(s_)) {
        w.mark = LastMark::PENDING;
    } else if (w.eager && w.done && hipEventRecord(w.done, s_) == hipSuccess) {
        w.mark = LastMark::RECORDED;
    } else if (w.eager && !w.done && hipEventCreateWithFlags(&w.done, hipEventDisableTiming) == hipSuccess &&
               hipEventRecord(w.done, s_) == hipSuccess) {
        w.mark = LastMark::RECORDED;
    } else {
        w.mark = LastMark::UNMARKED;
    }
    g.mu.unlock();
}

// ------------------------------------------------------------- InputFrames


struct InputFrames::State {
    Image spare;
    std::shared_ptr<StreamBox> copy;
    // per device buffer: its last readers, and the copy into it from the
    // pinned buffer of the same index (which frees that pinned buffer)
    hipEvent_t read[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    bool recorded[2] = {false, false};
    bool copying[2] = {false, false};
    void* pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
    int cur = 0;  // buffer the input image holds
    ~State() {
        for (int k = 0; k < 2; ++k) {
            if (copying[k]) (void)hipEventSynchronize(copied[k]);
            if (pinned[k]) (void)hipHostFree(pinned[k]);
            if (read[k]) (void)hipEventDestroy(read[k]);
            if (copied[k]) (void)hipEventDestroy(copied[k]);
        }
    }
};

InputFrames::InputFrames() = default;
InputFrames::InputFrames(InputFrames&&) noexcept = default;
InputFrames& InputFrames::operator=(InputFrames&&) noexcept = default;
InputFrames::~InputFrames() = default;

void InputFrames::load(Image& input, const ff_image_t& host) {
    if (!input.matches(host))
        throw invalid_argument("GPUImage::upload(): shapes do not match. Required: " + input.shapeString() +
                               ", passed: " + Image::shapeString(host));
    if (input.rowBytes() == 0 || input.height() == 0) return;
    if (!host.data) throw invalid_argument("GPUImage::upload(): null host data");
    if (host.pitch < input.rowBytes()) throw invalid_argument("GPUImage::upload(): host pitch smaller than a row");
    if (!st_) st_.reset(new State);
    State& t = *st_;
    const size_t pitch = input.pitch();
    const size_t span = (size_t)(input.height() - 1) * pitch + input.rowBytes();
    if (!t.spare.allocated() || !t.spare.sameShape(input) || t.spare.pitch() != pitch) {
        t.spare = Image(input.height(), input.width(), input.depth(), input.itemSize());
        t.recorded[0] = t.recorded[1] = false;
    }
    if (t.pinned_bytes < span) {
        for (int k = 0; k < 2; ++k) {
            if (t.copying[k]) FF_HIP(hipEventSynchronize(t.copied[k]));
            t.copying[k] = false;
            if (t.pinned[k]) FF_HIP(hipHostFree(t.pinned[k]));
            t.pinned[k] = nullptr;
        }
        t.pinned_bytes = 0;
        for (int k = 0; k < 2; ++k) FF_HIP(hipHostMalloc(&t.pinned[k], span, hipHostMallocDefault));
        t.pinned_bytes = span;
    }
    if (!t.copy) t.copy = StreamBox::create();
    for (int k = 0; k < 2; ++k) {
        if (!t.read[k]) FF_HIP(hipEventCreateWithFlags(&t.read[k], hipEventDisableTiming));
        if (!t.copied[k]) FF_HIP(hipEventCreateWithFlags(&t.copied[k], hipEventDisableTiming));
    }
    const int sp = t.cur ^ 1;
    hipStream_t cs = t.copy->get();
    // the pinned buffer is free once its last copy to the device is done
    if (t.copying[sp]) FF_HIP(hipEventSynchronize(t.copied[sp]));
    t.copying[sp] = false;
    HostCopier::get().copy((char*)t.pinned[sp], pitch, (const char*)host.data, host.pitch, input.rowBytes(),
                           input.height());
    // the device buffer is free once the work that read it is done
    if (t.recorded[sp]) FF_HIP(hipStreamWaitEvent(cs, t.read[sp], 0));
    FF_HIP(hipMemcpyAsync(t.spare.data(), t.pinned[sp], span, hipMemcpyHostToDevice, cs));
    FF_HIP(hipEventRecord(t.copied[sp], cs));
    t.copying[sp] = true;
    // everything enqueued later on the input's home stream reads the new frame
    FF_HIP(hipStreamWaitEvent(input.homeStream(), t.copied[sp], 0));
    input.swapBuffers(t.spare);
    t.cur = sp;
}

void InputFrames::readBy(hipStream_t s) {
    if (!st_ || !st_->read[st_->cur]) return;
    FF_HIP(hipEventRecord(st_->read[st_->cur], s));
    st_->recorded[st_->cur] = true;
}

void DeviceArray::assign(const void* host, size_t bytes) {
    st_ = std::make_shared<Storage>();
    if (bytes == 0) return;
    FF_HIP(hipMalloc(&st_->ptr, bytes));
    st_->bytes = bytes;
    hipStream_t s = stream();
    FF_HIP(hipMemcpyAsync(st_->ptr, host, bytes, hipMemcpyHostToDevice, s));
    FF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- profiler

namespace {
struct Profiler {
    std::mutex mu;
    bool on = false;
    std::string filter;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
};
Profiler& prof() {
    static Profiler p;
    return p;
}
}  // namespace

void profile_begin(const std::string& filter) {
    Profiler& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    p.on = true;
    p.filter = filter;
    p.used = 0;
}

void profile_end(double* total_ms, long long* count) {
    Profiler& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    double tot = 0.0;
    for (size_t i = 0; i < p.used; ++i) {
        FF_HIP(hipEventSynchronize(p.pool[i].second));
        float ms = 0.0f;
        FF_HIP(hipEventElapsedTime(&ms, p.pool[i].first, p.pool[i].second));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = (long long)p.used;
    p.on = false;
    p.used = 0;
}

bool profiling_active() {
    Profiler& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    return p.on;
}

ProfileScope::ProfileScope(const char* kernel, hipStream_t s) : s_(s) {
    Profiler& p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    if (!p.on || std::string(kernel).find(p.filter) == std::string::npos) return;
    if (p.used == p.pool.size()) {
        hipEvent_t a, b;
        FF_HIP(hipEventCreate(&a));
        FF_HIP(hipEventCreate(&b));
        p.pool.emplace_back(a, b);
    }
    slot_ = (int)p.used++;
    FF_HIP(hipEventRecord(p.pool[slot_].first, s_));
}

ProfileScope::~ProfileScope() {
    if (slot_ < 0) return;
    Profiler& p = prof();
    (void)hipEventRecord(p.pool[slot_].second, s_);
}

// ------------------------------------------------------------------- timer

Timer::~Timer() {
    if (start_) (void)hipEventDestroy(start_);
    if (stop_) (void)hipEventDestroy(stop_);
}

Timer& Timer::operator=(Timer&& o) noexcept {
    if (this != &o) {
        if (start_) (void)hipEventDestroy(start_);
        if (stop_) (void)hipEventDestroy(stop_);
        start_ = o.start_;
        stop_ = o.stop_;
        pending_ = o.pending_;
        last_ = o.last_;
        o.start_ = o.stop_ = nullptr;
        o.pending_ = false;
    }
    return *this;
}

void Timer::ensure() {
    if (!start_) FF_HIP(hipEventCreate(&start_));
    if (!stop_) FF_HIP(hipEventCreate(&stop_));
}

void Timer::start() {
    ensure();
    FF_HIP(hipEventRecord(start_, stream()));
}

void Timer::stop() {
    ensure();
    FF_HIP(hipEventRecord(stop_, stream()));
    pending_ = true;
}

// Stage::elapsedTime (src/gpu/pipeline.cu:87-89) after the event wait the
// reference performs in stopTiming (:72-82); here the wait happens only
// when the time is asked for.
float Timer::elapsed() {
    if (pending_) {
        FF_HIP(hipEventSynchronize(stop_));
        FF_HIP(hipEventElapsedTime(&last_, start_, stop_));
        pending_ = false;
    }
    return last_;
}

}  // namespace ffamd
