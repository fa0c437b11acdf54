// ff_filters.cpp -- see ff_filters.hpp. Reference: src/gpu/flowfilter.cu
// (line numbers per method).
#include "ff_filters.hpp"

#include <cmath>

namespace ffamd {

namespace {
std::string num(long v) { return std::to_string(v); }

void ensureLike(Image& spare, const Image& like) {
    if (!spare.allocated() || !spare.sameShape(like) || spare.pitch() != like.pitch())
        spare = Image(like.height(), like.width(), like.depth(), like.itemSize());
}
}  // namespace

// ================================================================ FlowFilter

// FlowFilter(GPUImage) (flowfilter.cu:32-38).
FlowFilter::FlowFilter(const Image& inputImage) {
    ownStream();
    StreamScope sc(box_);
    setInputImage(inputImage);
    configure();
}

// FlowFilter(h, w, smoothIterations, maxflow, gamma) (flowfilter.cu:45-76).
FlowFilter::FlowFilter(int height, int width, int smoothIterations, float maxflow, float gamma) {
    if (height <= 0)
        throw invalid_argument("FlowFilter::FlowFilter(): height should be greater than zero, got: " + num(height));
    if (width <= 0)
        throw invalid_argument("FlowFilter::FlowFilter(): width should be greater than zero, got: " + num(width));
    ownStream();
    StreamScope sc(box_);
    setInputImage(Image(height, width, 1, 1));
    ownInput_ = true;
    configure();
    setGamma(gamma);
    setMaxFlow(maxflow);
    setSmoothIterations(smoothIterations);
}

// FlowFilter::configure (flowfilter.cu:84-122): stages wired by shared
// buffers; the update's input flow is the propagator's output.
void FlowFilter::configure() {
    if (!inputSet_) throw logic_error("FlowFilter::configure(): input image has not been set");
    model_ = ImageModel(input_);
    Image dummyFlow(h_, w_, 2, 4);
    update_ = FlowUpdate(dummyFlow, model_.getImageConstant(), model_.getImageGradient(), 1.0f, 1.0f);
    smoother_ = FlowSmoother(update_.getUpdatedFlow(), 1);
    propagator_ = FlowPropagator(smoother_.getSmoothedFlow(), 1);
    update_.setInputFlow(propagator_.getPropagatedFlow());
    propagator_.getPropagatedFlow().clear();
    update_.getUpdatedFlow().clear();
    update_.getUpdatedImage().clear();
    smoother_.getSmoothedFlow().clear();
    spareFlow_ = Image();
    spareImage_ = Image();
    configured_ = true;
    firstLoad_ = true;
}

void FlowFilter::requireConfigured(const char* who) const {
    if (!configured_) throw logic_error(std::string("FlowFilter::") + who + ": filter not configured");
}

void FlowFilter::runImageModel() {
    requireConfigured("computeImageModel()");
    model_.run();
}

void FlowFilter::runPropagation() {
    requireConfigured("computePropagation()");
    propagator_.run();
}

// FlowFilter::computeUpdate (flowfilter.cu:175-197).
void FlowFilter::runUpdate() {
    requireConfigured("computeUpdate()");
    if (firstLoad_) {
        update_.getUpdatedImage().copyFrom(model_.getImageConstant());
        firstLoad_ = false;
    }
    update_.run();
    smoother_.run();
}

bool FlowFilter::fusedSupported() {
    if (!configured_ || firstLoad_) return false;
    return fused_level_supported(fusedArgs());
}

LevelArgs FlowFilter::fusedArgs() {
    Image S = smoother_.getSmoothedFlow();
    Image A = update_.getUpdatedImage();
    ensureLike(spareFlow_, S);
    ensureLike(spareImage_, A);
    LevelArgs a{};
    a.img = model_.boundImage().view();
    a.img_mode = model_.boundImage().itemSize() == 1 ? IMG_U8 : IMG_F32;
    a.flow_s_in = S.view();
    a.a0_in = A.view();
    a.flow_s_out = spareFlow_.view();
    a.a0_out = spareImage_.view();
    a.flow_upd_out = update_.getUpdatedFlow().view();
    a.gamma = update_.getGamma();
    a.maxflow = update_.getMaxFlow();
    a.prop_iters = propagator_.getIterations();
    a.dt = propagator_.getDt();
    a.border = propagator_.getBorder();
    a.smooth_iters = smoother_.getIterations();
    a.delta = 0;
    return a;
}

void FlowFilter::fusedCommit() {
    Image S = smoother_.getSmoothedFlow();
    Image A = update_.getUpdatedImage();
    S.swapBuffers(spareFlow_);
    A.swapBuffers(spareImage_);
}

// FlowFilter::compute (flowfilter.cu:125-153).
void FlowFilter::compute() {
    requireConfigured("compute()");
    if (model_.boundImage().height() != h_ || model_.boundImage().width() != w_)
        throw logic_error("FlowFilter::compute(): input image shape changed since configure()");
    startTiming();
    if (fused_ && fusedSupported()) {
        CrossStream x{&input_};
        const int t = frame_.prepare({fusedArgs()}, stream());
        frame_.launch(t, stream());
        check_launch("FlowFilter::compute (fused)");
        fusedCommit();
    } else {
        model_.run();
        if (firstLoad_) {
            update_.getUpdatedImage().copyFrom(model_.getImageConstant());
            firstLoad_ = false;
        }
        propagator_.run();
        update_.run();
        smoother_.run();
    }
    if (ownInput_) frames_.readBy(stream());
    stopTiming();
}

void FlowFilter::computeImageModel() {
    startTiming();
    runImageModel();
    if (ownInput_) frames_.readBy(stream());
    stopTiming();
}

void FlowFilter::computePropagation() {
    startTiming();
    runPropagation();
    stopTiming();
}

void FlowFilter::computeUpdate() {
    startTiming();
    runUpdate();
    stopTiming();
}

// FlowFilter::setInputImage (flowfilter.cu:200-216).
void FlowFilter::setInputImage(const Image& img) {
    if (img.depth() != 1)
        throw invalid_argument("FlowFilter::setInputImage(): input image should have depth 1, got: " +
                               num(img.depth()));
    if (img.itemSize() != 1 && img.itemSize() != 4)
        throw invalid_argument("FlowFilter::setInputImage(): item size should be 1 or 4: " + num(img.itemSize()));
    input_ = img;
    h_ = img.height();
    w_ = img.width();
    inputSet_ = true;
    ownInput_ = false;
    frames_ = InputFrames();
}

// FlowFilter::loadImage (flowfilter.cu:218-236). A filter that owns its
// input image loads host frames double buffered, overlapping the previous
// frame's compute().
void FlowFilter::loadImage(const ff_image_t& img) {
    if (ownInput_) frames_.load(input_, img);
    else input_.upload(img);
}
void FlowFilter::loadImageDevice(const void* dev, size_t pitch) { input_.upload_device(dev, pitch); }
void FlowFilter::downloadFlow(ff_image_t& flow) { smoother_.getSmoothedFlow().download(flow); }
void FlowFilter::downloadImage(ff_image_t& img) { update_.getUpdatedImage().download(img); }

// FlowFilter::setGamma (flowfilter.cu:276-285).
void FlowFilter::setGamma(float g) {
    if (input_.itemSize() == 1) update_.setGamma(g / (255.0f * 255.0f));
    else update_.setGamma(g);
}

// FlowFilter::setMaxFlow (flowfilter.cu:293-296).
void FlowFilter::setMaxFlow(float m) {
    update_.setMaxFlow(m);
    propagator_.setIterations(int(ceilf(m)));
}

// =========================================================== DeltaFlowFilter

// DeltaFlowFilter(inputImage, inputFlow) (flowfilter.cu:349-362).
DeltaFlowFilter::DeltaFlowFilter(const Image& inputImage, const Image& inputFlow) {
    ownStream();
    StreamScope sc(box_);
    setInputImage(inputImage);
    setInputFlow(inputFlow);
    configure();
}

// DeltaFlowFilter::configure (flowfilter.cu:371-426).
void DeltaFlowFilter::configure() {
    if (!flowSet_) throw runtime_error("DeltaFlowFilter::configure(): input flow not set");
    if (!imageSet_) throw runtime_error("DeltaFlowFilter::configure(): input image not set");
    const int h = input_.height(), w = input_.width();
    model_ = ImageModel(input_);
    Image dummyDelta(h, w, 2, 4), dummyOld(h, w, 1, 4);
    update_ = DeltaFlowUpdate(inputFlow_, dummyDelta, dummyOld, model_.getImageConstant(),
                              model_.getImageGradient());
    smoother_ = FlowSmoother(update_.getUpdatedFlow(), 1);
    propagator_ = FlowPropagatorPayload(smoother_.getSmoothedFlow(), update_.getUpdatedImage(),
                                        update_.getUpdatedDeltaFlow());
    update_.setInputDeltaFlow(propagator_.getPropagatedVector());
    update_.setInputImageOld(propagator_.getPropagatedScalar());
    model_.getImageConstant().clear();
    model_.getImageGradient().clear();
    propagator_.getPropagatedFlow().clear();
    propagator_.getPropagatedScalar().clear();
    propagator_.getPropagatedVector().clear();
    update_.getUpdatedFlow().clear();
    update_.getUpdatedDeltaFlow().clear();
    update_.getUpdatedImage().clear();
    smoother_.getSmoothedFlow().clear();
    spareFlow_ = Image();
    spareImage_ = Image();
    spareDelta_ = Image();
    configured_ = true;
    firstLoad_ = true;
}

void DeltaFlowFilter::requireConfigured(const char* who) const {
    if (!configured_) throw logic_error(std::string("DeltaFlowFilter::") + who + ": filter not configured");
}

void DeltaFlowFilter::runImageModel() {
    requireConfigured("computeImageModel()");
    model_.run();
}

void DeltaFlowFilter::runPropagation() {
    requireConfigured("computePropagation()");
    propagator_.run();
}

// DeltaFlowFilter::computeUpdate (flowfilter.cu:480-503).
void DeltaFlowFilter::runUpdate() {
    requireConfigured("computeUpdate()");
    if (firstLoad_) {
        update_.getUpdatedImage().copyFrom(model_.getImageConstant());
        propagator_.getPropagatedScalar().copyFrom(model_.getImageConstant());
        firstLoad_ = false;
    }
    update_.run();
    smoother_.run();
}

bool DeltaFlowFilter::fusedSupported() {
    if (!configured_ || firstLoad_) return false;
    return fused_level_supported(fusedArgs());
}

LevelArgs DeltaFlowFilter::fusedArgs() {
    Image S = smoother_.getSmoothedFlow();
    Image A = update_.getUpdatedImage();
    Image D = update_.getUpdatedDeltaFlow();
    ensureLike(spareFlow_, S);
    ensureLike(spareImage_, A);
    ensureLike(spareDelta_, D);
    LevelArgs a{};
    a.img = model_.boundImage().view();
    a.img_mode = model_.boundImage().itemSize() == 1 ? IMG_U8 : IMG_F32;
    a.flow_s_in = S.view();
    a.a0_in = A.view();
    a.dflow_in = D.view();
    a.flow_s_out = spareFlow_.view();
    a.a0_out = spareImage_.view();
    a.dflow_out = spareDelta_.view();
    a.flow_upd_out = update_.getUpdatedFlow().view();
    a.upper = update_.boundUpperFlow().view();
    a.xmap = update_.xmap();
    a.ymap = update_.ymap();
    a.gamma = update_.getGamma();
    a.maxflow = update_.getMaxFlow();
    a.prop_iters = propagator_.getIterations();
    a.dt = propagator_.getDt();
    a.border = propagator_.getBorder();
    a.smooth_iters = smoother_.getIterations();
    a.delta = 1;
    return a;
}

void DeltaFlowFilter::fusedCommit() {
    Image S = smoother_.getSmoothedFlow();
    Image A = update_.getUpdatedImage();
    Image D = update_.getUpdatedDeltaFlow();
    S.swapBuffers(spareFlow_);
    A.swapBuffers(spareImage_);
    D.swapBuffers(spareDelta_);
}

// DeltaFlowFilter::compute (flowfilter.cu:429-457). The first-load copy of
// this method precedes propagation (unlike computeUpdate's).
void DeltaFlowFilter::compute() {
    requireConfigured("compute()");
    startTiming();
    if (fused_ && fusedSupported()) {
        CrossStream x{&input_, &inputFlow_};
        const int t = frame_.prepare({fusedArgs()}, stream());
        frame_.launch(t, stream());
        check_launch("DeltaFlowFilter::compute (fused)");
        fusedCommit();
    } else {
        model_.run();
        if (firstLoad_) {
            update_.getUpdatedImage().copyFrom(model_.getImageConstant());
            firstLoad_ = false;
        }
        propagator_.run();
        update_.run();
        smoother_.run();
    }
    stopTiming();
}

void DeltaFlowFilter::computeImageModel() {
    startTiming();
    runImageModel();
    stopTiming();
}

void DeltaFlowFilter::computePropagation() {
    startTiming();
    runPropagation();
    stopTiming();
}

void DeltaFlowFilter::computeUpdate() {
    startTiming();
    runUpdate();
    stopTiming();
}

// DeltaFlowFilter::setInputImage / setInputFlow (flowfilter.cu:506-542).
void DeltaFlowFilter::setInputImage(const Image& img) {
    if (img.depth() != 1)
        throw runtime_error("DeltaFlowFilter::setInputImage(): input image should have depth 1: " + num(img.depth()));
    if (img.itemSize() != 1 && img.itemSize() != 4)
        throw runtime_error("DeltaFlowFilter::setInputImage(): input image should have item size 1 or 4: " +
                            num(img.itemSize()));
    input_ = img;
    imageSet_ = true;
}

void DeltaFlowFilter::setInputFlow(const Image& flow) {
    if (flow.depth() != 2)
        throw runtime_error("DeltaFlowFilter::setInputFlow(): input flow should have depth 2: " + num(flow.depth()));
    if (flow.itemSize() != 4)
        throw runtime_error("DeltaFlowFilter::setInputFlow(): input flow should have item size 4: " +
                            num(flow.itemSize()));
    inputFlow_ = flow;
    flowSet_ = true;
}

// DeltaFlowFilter::setGamma / setMaxFlow (flowfilter.cu:559-579).
void DeltaFlowFilter::setGamma(float g) {
    if (input_.itemSize() == 1) update_.setGamma(g / (255.0f * 255.0f));
    else update_.setGamma(g);
}

void DeltaFlowFilter::setMaxFlow(float m) {
    update_.setMaxFlow(m);
    propagator_.setIterations(int(ceilf(m)));
}

// ======================================================= PyramidalFlowFilter

// PyramidalFlowFilter(h, w, levels) (flowfilter.cu:631-641).
PyramidalFlowFilter::PyramidalFlowFilter(int height, int width, int levels)
    : h_(height), w_(width), levels_(levels) {
    ownStream();
    StreamScope sc(box_);
    configure();
}

// PyramidalFlowFilter::configure (flowfilter.cu:650-682): the top level is a
// FlowFilter on the coarsest image; each lower level a DeltaFlowFilter fed
// by the top's unsmoothed flow (first) or the next coarser level's
// smoothed flow.
void PyramidalFlowFilter::configure() {
    input_ = Image(h_, w_, 1, 1);
    pyramid_ = ImagePyramid(input_, levels_);
    top_ = FlowFilter(pyramid_.getImage(levels_ - 1));
    low_.clear();
    if (levels_ > 1) {
        low_.resize(levels_ - 1);
        Image levelInputFlow = top_.getFlow();
        levelInputFlow.clear();
        for (int l = levels_ - 2; l >= 0; --l) {
            low_[l] = DeltaFlowFilter(pyramid_.getImage(l), levelInputFlow);
            levelInputFlow = low_[l].getFlow();
        }
    }
    input_.clear();
    for (int l = 0; l < levels_; ++l) pyramid_.getImage(l).clear();
    setFused(fused_);
    graphs_.clear();
    configured_ = true;
}

void PyramidalFlowFilter::setFused(bool f) {
    fused_ = f;
    graphs_.clear();
    top_.setFused(f);
    for (auto& d : low_) d.setFused(f);
}

// PyramidalFlowFilter::compute (flowfilter.cu:685-716).
void PyramidalFlowFilter::computeStaged() {
    pyramid_.run();
    if (levels_ == 1) {
        // FlowFilter::compute (flowfilter.cu:125-153); its first-load copy
        // reads only the image model output, so it may follow propagation.
        top_.runImageModel();
        top_.runPropagation();
        top_.runUpdate();
        return;
    }
    top_.runImageModel();
    top_.runPropagation();
    for (int l = 0; l < levels_ - 1; ++l) {
        low_[l].runImageModel();
        low_[l].runPropagation();
    }
    top_.runUpdate();
    for (int l = 0; l < levels_ - 1; ++l) low_[l].runUpdate();
}

// ---------------------------------------------------------------- GraphCache

namespace {
bool same_img(const DevImg& a, const DevImg& b) {
    return a.data == b.data && a.h == b.h && a.w == b.w && a.pitch == b.pitch;
}
bool same_args(const LevelArgs& a, const LevelArgs& b) {
    return same_img(a.img, b.img) && a.img_mode == b.img_mode && a.wait_on == b.wait_on && same_img(a.flow_s_in, b.flow_s_in) &&
           same_img(a.a0_in, b.a0_in) && same_img(a.dflow_in, b.dflow_in) && same_img(a.flow_s_out, b.flow_s_out) &&
           same_img(a.a0_out, b.a0_out) && same_img(a.dflow_out, b.dflow_out) &&
           same_img(a.flow_upd_out, b.flow_upd_out) && same_img(a.upper, b.upper) && a.xmap == b.xmap &&
           a.ymap == b.ymap && a.gamma == b.gamma && a.maxflow == b.maxflow && a.dt == b.dt &&
           a.prop_iters == b.prop_iters && a.smooth_iters == b.smooth_iters && a.border == b.border &&
           a.delta == b.delta && a.frame_slot == b.frame_slot;
}
}  // namespace

GraphCache& GraphCache::operator=(GraphCache&& o) noexcept {
    if (this != &o) {
        clear();
        for (int i = 0; i < 2; ++i) {
            slot_[i] = std::move(o.slot_[i]);
            o.slot_[i].exec = nullptr;
            o.slot_[i].key.clear();
            o.slot_[i].token = 0;
        }
        next_ = o.next_;
    }
    return *this;
}

hipGraphExec_t GraphCache::find(const std::vector<LevelArgs>& key, unsigned long long token) const {
    for (const Entry& e : slot_) {
        if (!e.exec || e.token != token || e.key.size() != key.size()) continue;
        bool eq = true;
        for (size_t i = 0; i < key.size() && eq; ++i) eq = same_args(e.key[i], key[i]);
        if (eq) return e.exec;
    }
    return nullptr;
}

void GraphCache::insert(const std::vector<LevelArgs>& key, unsigned long long token, hipGraphExec_t exec) {
    Entry& e = slot_[next_];
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.exec = exec;
    e.key = key;
    e.token = token;
    next_ ^= 1;
}

void GraphCache::clear() {
    for (Entry& e : slot_) {
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
        e.exec = nullptr;
        e.key.clear();
        e.token = 0;
    }
}

// The same frame as one fused launch. Jobs: the top level first, then
// levels 0 .. L-2. Level l reads level l+1's smoothed flow before level l+1
// has advanced (the reference's update order); level L-2 reads the top
// level's unsmoothed flow of this frame and waits for it inside the launch.
// Every level writes its new state to its spare buffers, exchanged with the
// state after the launch.
bool PyramidalFlowFilter::fusedPlan(std::vector<LevelArgs>& args) {
    args.clear();
    if (!configured_ || top_.firstLoad()) return false;
    for (const DeltaFlowFilter& d : low_)
        if (d.firstLoad()) return false;
    // Levels 1 and 2 reduce the input image while loading it (one or two
    // pyramid steps), so with up to three levels the pyramid kernels do not
    // run; deeper pyramids read the pyramid's images.
    const bool reduce = levels_ <= 3;
    LevelArgs top = top_.fusedArgs();
    if (reduce && levels_ > 1 && top.img_mode == IMG_U8) {
        top.img = input_.view();
        top.img_mode = levels_ == 2 ? IMG_U8_DOWN : IMG_U8_DOWN2;
    }
    args.push_back(top);
    for (int l = 0; l < levels_ - 1; ++l) {
        LevelArgs a = low_[l].fusedArgs();
        if (reduce && l == 1 && a.img_mode == IMG_U8) {
            a.img = input_.view();
            a.img_mode = IMG_U8_DOWN;
        }
        if (l == levels_ - 2) a.wait_on = 0;
        args.push_back(a);
    }
    for (LevelArgs& a : args) {
        if (!fused_level_supported(a)) return false;
        // levels that read the input frame itself may take it in place
        a.frame_slot = reduce && a.img.data == input_.view().data ? 0 : -1;
    }
    return true;
}

void PyramidalFlowFilter::enqueuePreFused() {
    if (levels_ > 3) pyramid_.run();
}

void PyramidalFlowFilter::commitFused() {
    top_.fusedCommit();
    for (int l = 0; l < levels_ - 1; ++l) low_[l].fusedCommit();
}

void PyramidalFlowFilter::enqueueFused(int table, hipStream_t s, const FrameRef* frame) {
    enqueuePreFused();
    frame_.launch(table, s, frame, frame ? 1 : 0);
    check_launch("PyramidalFlowFilter::compute (fused frame)");
}

namespace {
// Enqueue `work` on s, through a graph of `cache` keyed by (args, token)
// when graphs are on and no kernel profile is open (profiling brackets each
// launch with events, which a replay would not do). Returns whether a graph
// was launched.
template <class F>
bool run_maybe_graphed(bool graphs_on, GraphCache& cache, const std::vector<LevelArgs>& args,
                       unsigned long long token, hipStream_t s, F work) {
    if (!graphs_on || profiling_active()) {
        work();
        return false;
    }
    hipGraphExec_t exec = cache.find(args, token);
    if (!exec) {
        FF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
            work();
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        FF_HIP(hipStreamEndCapture(s, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw_hip(e, "hipGraphInstantiate", __FILE__, __LINE__);
        cache.insert(args, token, exec);
    }
    FF_HIP(hipGraphLaunch(exec, s));
    return true;
}
}  // namespace

void PyramidalFlowFilter::computeFused(const std::vector<LevelArgs>& args, const void* frame, size_t frame_pitch) {
    hipStream_t s = stream();
    const int table = frame_.prepare(args, s);
    // a graph pays off only when the frame is more than the one launch
    // (pyramid kernels for more than three levels); replaying a one-node
    // graph costs more than launching the kernel
    const FrameRef fr{frame, frame_pitch};
    const WaitingLaunch order(s, frame_.waits(table));
    // a frame in place is a launch argument, which a replayed graph would
    // not update: such frames are launched directly
    if (run_maybe_graphed(graphs_on_ && levels_ > 3 && !frame, graphs_, args, frame_.token(table), s,
                          [&] { enqueueFused(table, s, frame ? &fr : nullptr); }))
        frame_.replayed();
    commitFused();
}

bool PyramidalFlowFilter::frameFits(const void* dev, size_t pitch) const {
    return dev && (uintptr_t)dev % 4 == 0 && pitch % 4 == 0 && (unsigned long long)pitch * (unsigned long long)h_ < 0x7fffffffULL;
}

void PyramidalFlowFilter::compute() {
    if (!configured_) throw logic_error("PyramidalFlowFilter::compute(): filter not configured");
    startTiming();
    const void* frame = frame_dev_;
    const size_t pitch = frame_pitch_;
    frame_dev_ = nullptr;
    std::vector<LevelArgs> args;
    const bool fused = fused_ && fusedPlan(args);
    // a frame passed in place is read by the fused launch (up to three
    // levels: every level reduces the input frame while loading it), else
    // copied into the input image first
    const bool in_place = frame && fused && levels_ <= 3 && frameFits(frame, pitch);
    if (frame && !in_place) input_.upload_device(frame, pitch);
    if (fused) computeFused(args, in_place ? frame : nullptr, in_place ? pitch : 0);
    else computeStaged();
    frames_.readBy(stream());
    stopTiming();
}

void PyramidalFlowFilter::checkWaits() {
    waitOwn_.follow(frame_.watch());
    waitOwn_.check("PyramidalFlowFilter");
    waitBatch_.check("PyramidalFlowFilter (in a batch)");
}

float PyramidalFlowFilter::elapsedTime() {
    const float t = Stage::elapsedTime();
    checkWaits();
    return t;
}

Image PyramidalFlowFilter::getFlow() const { return levels_ == 1 ? top_.getFlow() : low_[0].getFlow(); }

// PyramidalFlowFilter::loadImage (flowfilter.cu:728-731): double buffered
// (InputFrames), so the copy overlaps the previous frame's compute().
void PyramidalFlowFilter::loadImage(const ff_image_t& img) {
    frames_.load(input_, img);
    frame_dev_ = nullptr;
}
void PyramidalFlowFilter::loadImageDevice(const void* dev, size_t pitch) {
    input_.upload_device(dev, pitch);
    frame_dev_ = nullptr;
}
void PyramidalFlowFilter::setImageDevice(const void* dev, size_t pitch) {
    if (!dev) throw invalid_argument("PyramidalFlowFilter::setImageDevice(): null frame");
    if (!configured_) throw logic_error("PyramidalFlowFilter::setImageDevice(): filter not configured");
    if (pitch < (size_t)w_)
        throw invalid_argument("PyramidalFlowFilter::setImageDevice(): pitch " + std::to_string(pitch) +
                               " is less than the width " + std::to_string(w_));
    frame_dev_ = dev;
    frame_pitch_ = pitch;
}

void PyramidalFlowFilter::downloadFlow(ff_image_t& flow) {
    if (levels_ == 1) top_.downloadFlow(flow);
    else low_[0].getFlow().download(flow);
    checkWaits();
}

void PyramidalFlowFilter::downloadImage(ff_image_t& img) {
    if (levels_ == 1) top_.downloadImage(img);
    else low_[0].getImage().download(img);
    checkWaits();
}

void PyramidalFlowFilter::checkLevel(int level, const char* who) const {
    if (level < 0 || level >= levels_)
        throw runtime_error(std::string("PyramidalFlowFilter::") + who + ": level index out of bounds: " + num(level));
}

float PyramidalFlowFilter::getGamma(int level) const {
    checkLevel(level, "getGamma()");
    return level == levels_ - 1 ? top_.getGamma() : low_[level].getGamma();
}

void PyramidalFlowFilter::setGamma(int level, float g) {
    checkLevel(level, "setGamma()");
    if (level == levels_ - 1) top_.setGamma(g);
    else low_[level].setGamma(g);
}

void PyramidalFlowFilter::setGamma(const std::vector<float>& g) {
    if ((int)g.size() != levels_)
        throw runtime_error("PyramidalFlowFilter::setGamma(): gamma vector should be size " + num(levels_) +
                            ", got: " + num((long)g.size()));
    for (int l = 0; l < levels_; ++l) setGamma(l, g[l]);
}

int PyramidalFlowFilter::getSmoothIterations(int level) const {
    checkLevel(level, "getSmoothIterations()");
    return level == levels_ - 1 ? top_.getSmoothIterations() : low_[level].getSmoothIterations();
}

void PyramidalFlowFilter::setSmoothIterations(int level, int n) {
    checkLevel(level, "setSmoothIterations()");
    if (level == levels_ - 1) top_.setSmoothIterations(n);
    else low_[level].setSmoothIterations(n);
}

void PyramidalFlowFilter::setSmoothIterations(const std::vector<int>& n) {
    if ((int)n.size() != levels_)
        throw runtime_error("PyramidalFlowFilter::setSmoothIterations(): iterations vector should be size " +
                            num(levels_) + ", got: " + num((long)n.size()));
    for (int l = 0; l < levels_; ++l) setSmoothIterations(l, n[l]);
}

float PyramidalFlowFilter::getMaxFlow() const { return levels_ == 1 ? top_.getMaxFlow() : low_[0].getMaxFlow(); }

// PyramidalFlowFilter::setMaxFlow (flowfilter.cu:849-866): level 0 gets m,
// each coarser level half of the previous, the top the last halving.
void PyramidalFlowFilter::setMaxFlow(float m) {
    if (levels_ == 1) {
        top_.setMaxFlow(m);
        return;
    }
    float ml = m;
    for (int l = 0; l < levels_ - 1; ++l) {
        low_[l].setMaxFlow(ml);
        ml /= 2.0f;
    }
    top_.setMaxFlow(ml);
}

// PyramidalFlowFilter::setPropagationBorder (flowfilter.cu:869-877). The
// reference loops h < levels over levels-1 filters (reading one past the
// end); here every existing level is set.
void PyramidalFlowFilter::setPropagationBorder(int b) {
    top_.setPropagationBorder(b);
    for (auto& d : low_) d.setPropagationBorder(b);
}

// ------------------------------------------------------- BatchedFlowFilter

// A filter in a batch enqueues all its work on the batch's stream (its own
// compute() and transfers too), so the batched launch, which reads and
// writes its state, is ordered with them.
int BatchedFlowFilter::add(const std::shared_ptr<PyramidalFlowFilter>& f) {
    if (!f) throw invalid_argument("BatchedFlowFilter::add(): null filter");
    for (const auto& g : streams_)
        if (g == f) throw invalid_argument("BatchedFlowFilter::add(): filter already in the batch");
    f->box()->follow(box_);
    streams_.push_back(f);
    graphs_.clear();
    return (int)streams_.size() - 1;
}

void BatchedFlowFilter::setImagesDevice(const void* const* ptrs, const size_t* pitches, int n) {
    if (n != (int)streams_.size())
        throw invalid_argument("BatchedFlowFilter::setImagesDevice(): " + num(n) + " frames for " +
                               num((long)streams_.size()) + " streams");
    if (n > 0 && (!ptrs || !pitches)) throw invalid_argument("BatchedFlowFilter::setImagesDevice(): null array");
    for (int i = 0; i < n; ++i) streams_[i]->setImageDevice(ptrs[i], pitches[i]);
}

float BatchedFlowFilter::elapsedTime() {
    const float t = Stage::elapsedTime();
    waitOwn_.follow(frame_.watch());
    waitOwn_.check("BatchedFlowFilter");
    return t;
}

void BatchedFlowFilter::loadImagesDevice(const void* const* ptrs, const size_t* pitches, int n) {
    if (n != (int)streams_.size())
        throw invalid_argument("BatchedFlowFilter::loadImagesDevice(): " + num(n) + " frames for " +
                               num((long)streams_.size()) + " streams");
    if (n > 0 && (!ptrs || !pitches)) throw invalid_argument("BatchedFlowFilter::loadImagesDevice(): null array");
    std::vector<CopyRegion> r;
    r.reserve(n);
    for (int i = 0; i < n; ++i) {
        Image& in = streams_[i]->input_;
        if (!ptrs[i]) throw invalid_argument("BatchedFlowFilter::loadImagesDevice(): null frame " + num(i));
        if (pitches[i] < in.rowBytes())
            throw invalid_argument("BatchedFlowFilter::loadImagesDevice(): pitch of frame " + num(i) +
                                   " smaller than a row");
        r.push_back(CopyRegion{ptrs[i], in.data(), pitches[i], in.pitch(), in.rowBytes(), in.height()});
    }
    launch_copy_regions(r.data(), n, ffamd::stream());
    check_launch("BatchedFlowFilter::loadImagesDevice");
}

const std::shared_ptr<PyramidalFlowFilter>& BatchedFlowFilter::stream(int i) const {
    if (i < 0 || i >= (int)streams_.size())
        throw invalid_argument("BatchedFlowFilter::stream(): index out of range: " + num(i));
    return streams_[i];
}

// One frame step of every stream: the jobs of all streams whose next frame
// can run fused go into one launch (their level jobs keep their order and
// their wait on their own top level); the others run their stage kernels.
void BatchedFlowFilter::compute() {
    if (streams_.empty()) throw logic_error("BatchedFlowFilter::compute(): no streams");
    startTiming();
    std::vector<LevelArgs> all, one;
    std::vector<PyramidalFlowFilter*> fused;
    std::vector<FrameRef> frames;
    for (const auto& f : streams_) {
        // a filter moved to another stream since it joined comes back
        if (f->box()->get() != box_->get()) f->box()->follow(box_);
        const void* fp = f->frame_dev_;
        const size_t pp = f->frame_pitch_;
        f->frame_dev_ = nullptr;
        const bool fusedOk = f->fused() && f->fusedPlan(one);
        // a frame given in place is read by the launch where it lies (one
        // launch argument per stream), else copied into the input image
        const bool in_place = fp && fusedOk && f->levels_ <= 3 && f->frameFits(fp, pp) &&
                              (int)frames.size() < kMaxLaunchFrames;
        if (fp && !in_place) f->input_.upload_device(fp, pp);
        if (fusedOk) {
            const int slot = in_place ? (int)frames.size() : -1;
            if (in_place) frames.push_back(FrameRef{fp, pp});
            const int off = (int)all.size();
            for (LevelArgs a : one) {
                if (a.wait_on >= 0) a.wait_on += off;
                a.frame_slot = a.frame_slot >= 0 ? slot : -1;
                all.push_back(a);
            }
            fused.push_back(f.get());
        } else {
            f->computeStaged();
        }
    }
    if (!all.empty()) {
        hipStream_t s = ffamd::stream();
        const int table = frame_.prepare(all, s);
        last_blocks_ = frame_.blocks(table);
        const WaitingLaunch order(s, frame_.waits(table));
        // the step is replayed as a HIP graph when graphs are on (the
        // default) and no frame is read in place (a launch argument, which
        // a replay would not update)
        if (run_maybe_graphed(graphs_on_ && frames.empty(), graphs_, all, frame_.token(table), s, [&] {
                for (PyramidalFlowFilter* f : fused) f->enqueuePreFused();
                frame_.launch(table, s, frames.data(), (int)frames.size());
                check_launch("BatchedFlowFilter::compute (fused frame)");
            }))
            frame_.replayed();
        for (PyramidalFlowFilter* f : fused) {
            f->commitFused();
            f->waitBatch_.follow(frame_.watch());
        }
    }
    for (const auto& f : streams_) f->frames_.readBy(ffamd::stream());
    stopTiming();
}

}  // namespace ffamd
