// ff_filters.hpp -- the filter classes of flowfilter::gpu
// (include/flowfilter/gpu/flowfilter.h:31-332): FlowFilter,
// DeltaFlowFilter and PyramidalFlowFilter.
//
// Two execution paths produce identical results:
//  - staged: one launch per reference stage kernel, the reference's exact
//    sequence (used for the first frame, for the split compute methods and
//    when the fused kernels do not cover a parameter set);
//  - fused: one kernel launch per frame for all levels (image model,
//    propagation, update and smoothing of each level in one pass over HBM),
//    with the level state ping-ponged between two buffers.
#pragma once

#include <memory>
#include <vector>

#include "ff_stages.hpp"

namespace ffamd {

class FlowFilter : public Stage {
public:
    FlowFilter() = default;
    explicit FlowFilter(const Image& inputImage);
    FlowFilter(int height, int width, int smoothIterations = 1, float maxflow = 1.0f, float gamma = 1.0f);
    void configure() override;
    void compute() override;
    const char* name() const override { return "FlowFilter"; }
    void computeImageModel();
    void computePropagation();
    void computeUpdate();
    void setInputImage(const Image& img);
    Image getFlow() const { return update_.getUpdatedFlow(); }
    void loadImage(const ff_image_t& img);
    void loadImageDevice(const void* dev, size_t pitch);
    void downloadFlow(ff_image_t& flow);
    void downloadImage(ff_image_t& img);
    float getGamma() const { return update_.getGamma(); }
    void setGamma(float g);
    float getMaxFlow() const { return update_.getMaxFlow(); }
    void setMaxFlow(float m);
    int getSmoothIterations() const { return smoother_.getIterations(); }
    void setSmoothIterations(int n) { smoother_.setIterations(n); }
    void setPropagationBorder(int b) { propagator_.setBorder(b); }
    int getPropagationBorder() const { return propagator_.getBorder(); }
    int getPropagationIterations() const { return propagator_.getIterations(); }
    int height() const { return h_; }
    int width() const { return w_; }
    void setFused(bool f) { fused_ = f; }
    bool fused() const { return fused_; }
    void setTileWaves(int nw) { frame_.setWaves(nw); }
    int tileWaves() const { return frame_.waves(); }
    void setTileTopRows(int rows) { frame_.setTopRows(rows); }
    int tileTopRows() const { return frame_.topRows(); }
    int tileRows() const { return frame_.rows(); }
    void setTileRows(int rows) { frame_.setRows(rows); }
    int tileInterior() const { return frame_.interior(); }
    void setTileInterior(int mode) { frame_.setInterior(mode); }

    // Used by PyramidalFlowFilter: the same work without the timer.
    void runImageModel();
    void runPropagation();
    void runUpdate();
    bool firstLoad() const { return firstLoad_; }
    // Arguments of the fused kernel for this level (allocates the spare
    // state buffers on first use) and the state exchange after it ran.
    LevelArgs fusedArgs();
    void fusedCommit();
    bool fusedSupported();

private:
    void requireConfigured(const char* who) const;
    int h_ = 0, w_ = 0;
    bool configured_ = false, firstLoad_ = false, inputSet_ = false, fused_ = true;
    // the input image is the filter's own (FlowFilter(h, w, ...)): host
    // frames load double buffered (InputFrames)
    bool ownInput_ = false;
    Image input_;
    InputFrames frames_;
    ImageModel model_;
    FlowUpdate update_;
    FlowSmoother smoother_;
    FlowPropagator propagator_;
    Image spareFlow_, spareImage_;
    FusedFrame frame_;
};

class DeltaFlowFilter : public Stage {
public:
    DeltaFlowFilter() = default;
    DeltaFlowFilter(const Image& inputImage, const Image& inputFlow);
    void configure() override;
    void compute() override;
    const char* name() const override { return "DeltaFlowFilter"; }
    void computeImageModel();
    void computePropagation();
    void computeUpdate();
    void setInputImage(const Image& img);
    void setInputFlow(const Image& flow);
    Image getFlow() const { return smoother_.getSmoothedFlow(); }
    Image getImage() const { return update_.getUpdatedImage(); }
    float getGamma() const { return update_.getGamma(); }
    void setGamma(float g);
    float getMaxFlow() const { return update_.getMaxFlow(); }
    void setMaxFlow(float m);
    int getSmoothIterations() const { return smoother_.getIterations(); }
    void setSmoothIterations(int n) { smoother_.setIterations(n); }
    void setPropagationBorder(int b) { propagator_.setBorder(b); }
    int getPropagationBorder() const { return propagator_.getBorder(); }
    int getPropagationIterations() const { return propagator_.getIterations(); }
    int height() const { return input_.height(); }
    int width() const { return input_.width(); }
    void setFused(bool f) { fused_ = f; }
    bool fused() const { return fused_; }
    void setTileWaves(int nw) { frame_.setWaves(nw); }
    int tileWaves() const { return frame_.waves(); }
    void setTileTopRows(int rows) { frame_.setTopRows(rows); }
    int tileTopRows() const { return frame_.topRows(); }
    int tileRows() const { return frame_.rows(); }
    void setTileRows(int rows) { frame_.setRows(rows); }
    int tileInterior() const { return frame_.interior(); }
    void setTileInterior(int mode) { frame_.setInterior(mode); }

    void runImageModel();
    void runPropagation();
    void runUpdate();
    bool firstLoad() const { return firstLoad_; }
    LevelArgs fusedArgs();
    void fusedCommit();
    bool fusedSupported();

private:
    void requireConfigured(const char* who) const;
    bool configured_ = false, firstLoad_ = true, imageSet_ = false, flowSet_ = false, fused_ = true;
    Image input_, inputFlow_;
    ImageModel model_;
    DeltaFlowUpdate update_;
    FlowSmoother smoother_;
    FlowPropagatorPayload propagator_;
    Image spareFlow_, spareImage_, spareDelta_;
    FusedFrame frame_;
};

// Instantiated HIP graphs of one frame's fused launch sequence, keyed by the
// exact level arguments and the identity of the job table's device memory.
// The level state alternates between two buffers, so a steady run needs two
// graphs (even and odd frames); any parameter or buffer change produces a
// new key and a new capture.
class GraphCache {
public:
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    GraphCache(GraphCache&& o) noexcept { *this = std::move(o); }
    GraphCache& operator=(GraphCache&& o) noexcept;
    ~GraphCache() { clear(); }
    hipGraphExec_t find(const std::vector<LevelArgs>& key, unsigned long long token) const;
    void insert(const std::vector<LevelArgs>& key, unsigned long long token, hipGraphExec_t exec);
    void clear();

private:
    struct Entry {
        std::vector<LevelArgs> key;
        unsigned long long token = 0;
        hipGraphExec_t exec = nullptr;
    };
    Entry slot_[2];
    int next_ = 0;
};

class BatchedFlowFilter;

class PyramidalFlowFilter : public Stage {
public:
    PyramidalFlowFilter() = default;
    PyramidalFlowFilter(int height, int width, int levels);
    void configure() override;
    void compute() override;
    const char* name() const override { return "PyramidalFlowFilter"; }
    Image getFlow() const;
    void loadImage(const ff_image_t& img);
    void loadImageDevice(const void* dev, size_t pitch);
    // The next compute() reads its frame where the caller keeps it (device
    // u8, the filter's height x width, row pitch `pitch` bytes): the fused
    // frame kernel loads it directly instead of a copy into the input image.
    // The memory must hold the frame until that compute() has run on the
    // stream. Falls back to a copy (loadImageDevice) for the first frame,
    // pyramids deeper than three levels, the staged path, or a pointer or
    // pitch that is not a multiple of 4 bytes. No reference counterpart.
    void setImageDevice(const void* dev, size_t pitch);
    void downloadFlow(ff_image_t& flow);
    void downloadImage(ff_image_t& img);
    // Waits for the last compute(); FF_ERR_RUNTIME when a wait of the fused
    // kernel gave up since the last synchronizing call (downloads too).
    float elapsedTime() override;
    float getGamma(int level) const;
    void setGamma(int level, float g);
    void setGamma(const std::vector<float>& g);
    int getSmoothIterations(int level) const;
    void setSmoothIterations(int level, int n);
    void setSmoothIterations(const std::vector<int>& n);
    float getMaxFlow() const;
    void setMaxFlow(float m);
    void setPropagationBorder(int b);
    int getPropagationBorder() const { return top_.getPropagationBorder(); }
    int height() const { return h_; }
    int width() const { return w_; }
    int levels() const { return levels_; }
    void setFused(bool f);
    bool fused() const { return fused_; }
    // Replay the fused frame as a captured HIP graph (default on).
    void setGraphs(bool g) {
        graphs_on_ = g;
        graphs_.clear();
    }
    bool graphs() const { return graphs_on_; }
    unsigned long long waitTimeouts() const { return frame_.timeouts(); }
    // Test hook (FusedFrame::setWaitTest).
    void setWaitTest(int spins) {
        frame_.setWaitTest(spins);
        graphs_.clear();
    }
    // Waves per workgroup of the fused frame kernel: 0 = autotune.
    void setTileWaves(int nw) { frame_.setWaves(nw); }
    int tileWaves() const { return frame_.waves(); }
    void setTileTopRows(int rows) { frame_.setTopRows(rows); }
    int tileTopRows() const { return frame_.topRows(); }
    int tileRows() const { return frame_.rows(); }
    void setTileRows(int rows) { frame_.setRows(rows); }
    int tileInterior() const { return frame_.interior(); }
    void setTileInterior(int mode) { frame_.setInterior(mode); }

private:
    friend class BatchedFlowFilter;
    void checkLevel(int level, const char* who) const;
    // FF_ERR_RUNTIME when a wait of this filter's fused launches (or of a
    // batch's that advanced it) gave up since the last check.
    void checkWaits();
    // Whether a frame at (dev, pitch) can be read in place by the fused kernel.
    bool frameFits(const void* dev, size_t pitch) const;
    void computeStaged();
    // Pyramid kernels a fused frame needs before its launch (deeper pyramids)
    // and the state exchange after it.
    void enqueuePreFused();
    void commitFused();
    // Level arguments of a fused frame; false when some level is not ready
    // (first frame) or not covered by the fused kernel.
    bool fusedPlan(std::vector<LevelArgs>& args);
    void computeFused(const std::vector<LevelArgs>& args, const void* frame = nullptr, size_t frame_pitch = 0);
    void enqueueFused(int table, hipStream_t s, const FrameRef* frame = nullptr);
    bool configured_ = false, fused_ = true, graphs_on_ = true;
    // frame passed in place for the next compute() (setImageDevice)
    const void* frame_dev_ = nullptr;
    size_t frame_pitch_ = 0;
    GraphCache graphs_;
    FusedFrame frame_;
    WaitCheck waitOwn_, waitBatch_;
    int h_ = 0, w_ = 0, levels_ = 0;
    Image input_;
    InputFrames frames_;
    ImagePyramid pyramid_;
    FlowFilter top_;
    std::vector<DeltaFlowFilter> low_;
};

// Many PyramidalFlowFilters (camera streams of any size and depth) advanced
// together: compute() moves every stream one frame forward with a single
// fused frame launch (replayed as a HIP graph); each stream's results are
// identical to its own compute(). Streams in their first frame run their
// stage kernels first. No reference counterpart.
class BatchedFlowFilter : public Stage {
public:
    BatchedFlowFilter() { ownStream(); }
    void configure() override {}
    void compute() override;
    const char* name() const override { return "BatchedFlowFilter"; }
    int add(const std::shared_ptr<PyramidalFlowFilter>& f);
    int size() const { return (int)streams_.size(); }
    // Load the next frame of every stream from device memory (u8, pitched)
    // in one copy launch; n must equal size().
    void loadImagesDevice(const void* const* ptrs, const size_t* pitches, int n);
    // The next compute() reads every stream's frame where the caller keeps
    // it (setImageDevice of each stream; the fused launch takes them in
    // place, up to kMaxLaunchFrames of them).
    void setImagesDevice(const void* const* ptrs, const size_t* pitches, int n);
    float elapsedTime() override;
    const std::shared_ptr<PyramidalFlowFilter>& stream(int i) const;
    void setGraphs(bool g) {
        graphs_on_ = g;
        graphs_.clear();
    }
    bool graphs() const { return graphs_on_; }
    unsigned long long waitTimeouts() const { return frame_.timeouts(); }
    void setWaitTest(int spins) {
        frame_.setWaitTest(spins);
        graphs_.clear();
    }
    void setTileWaves(int nw) { frame_.setWaves(nw); }
    int tileWaves() const { return frame_.waves(); }
    void setTileTopRows(int rows) { frame_.setTopRows(rows); }
    int tileTopRows() const { return frame_.topRows(); }
    int tileRows() const { return frame_.rows(); }
    void setTileRows(int rows) { frame_.setRows(rows); }
    int tileInterior() const { return frame_.interior(); }
    void setTileInterior(int mode) { frame_.setInterior(mode); }
    // Workgroups of the last fused launch (0 before the first).
    int lastBlocks() const { return last_blocks_; }

private:
    std::vector<std::shared_ptr<PyramidalFlowFilter>> streams_;
    bool graphs_on_ = true;
    int last_blocks_ = 0;
    GraphCache graphs_;
    FusedFrame frame_;
    WaitCheck waitOwn_;
};

}  // namespace ffamd
