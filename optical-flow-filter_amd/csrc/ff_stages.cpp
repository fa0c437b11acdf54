// ff_stages.cpp -- see ff_stages.hpp. Reference line numbers are given per
// method (src/gpu/<file>.cu).
#include "ff_stages.hpp"

#include <cmath>
#include <sstream>

namespace ffamd {

namespace {

std::string num(long v) { return std::to_string(v); }

void require_shape(const Image& a, const Image& b, const char* who) {
    if (a.height() != b.height() || a.width() != b.width())
        throw logic_error(std::string(who) + ": input buffers do not match height and width " + a.shapeString() +
                          " vs " + b.shapeString());
}

void require_alloc(const Image& a, const char* who) {
    if (!a.allocated() && (size_t)a.height() * a.rowBytes() > 0)
        throw logic_error(std::string(who) + ": unallocated buffer");
}

}  // namespace

std::vector<int> upper_index_map(int n, int nu) {
    std::vector<int> m((size_t)(n > 0 ? n : 0), 0);
    for (int x = 0; x < n; ++x) {
        int i = 0;
        if (n > 1 && nu > 0) {
            const float u = (float)x / (float)(n - 1);
            const float s = u * (float)nu;
            i = (int)std::floor(s);
            if (i > nu - 1) i = nu - 1;
            if (i < 0) i = 0;
        }
        m[x] = i;
    }
    return m;
}

// ================================================================ ImageModel
// src/gpu/imagemodel.cu:24-146

ImageModel::ImageModel(const Image& input) {
    setInputImage(input);
    configure();
}

void ImageModel::configure() {
    if (!inputSet_) throw logic_error("ImageModel::configure(): input image has not been set");
    bound_ = input_;
    const int h = bound_.height(), w = bound_.width();
    a0_ = Image(h, w, 1, 4);
    grad_ = Image(h, w, 2, 4);
    configured_ = true;
}

void ImageModel::setInputImage(const Image& img) {
    if (img.depth() != 1)
        throw invalid_argument("ImageModel::setInputImage(): image depth should be 1, got: " + num(img.depth()));
    if (img.itemSize() != 1 && img.itemSize() != 4)
        throw invalid_argument("ImageModel::setInputImage(): item size should be 1 or 4, got: " + num(img.itemSize()));
    input_ = img;
    inputSet_ = true;
}

void ImageModel::run() {
    if (!configured_) throw logic_error("ImageModel::compute() stage not configured.");
    require_shape(bound_, a0_, "ImageModel::compute()");
    require_alloc(bound_, "ImageModel::compute()");
    CrossStream x{&bound_, &a0_, &grad_};
    launch_image_model(bound_.view(), bound_.itemSize() == 1, a0_.view(), grad_.view(), stream());
    check_launch("ImageModel::compute");
}

void ImageModel::compute() {
    startTiming();
    run();
    stopTiming();
}

// ============================================================ FlowPropagator
// src/gpu/propagation.cu:20-211

FlowPropagator::FlowPropagator(const Image& inputFlow, int iterations) {
    setInputFlow(inputFlow);
    setIterations(iterations);
    configure();
}

void FlowPropagator::configure() {
    if (!inputSet_) throw runtime_error("FlowPropagator::configure(): input flow has not been set");
    bound_ = input_;
    bufX_ = Image(bound_.height(), bound_.width(), 2, 4);
    bufY_ = Image(bound_.height(), bound_.width(), 2, 4);
    configured_ = true;
}

void FlowPropagator::setIterations(int n) {
    if (n <= 0) throw runtime_error("FlowPropagator::setIterations(): iterations less than zero: " + num(n));
    iterations_ = n;
    dt_ = 1.0f / float(iterations_);
}

void FlowPropagator::setBorder(int b) {
    if (b < 0) throw runtime_error("FlowPropagator::setBorder(): border should be greater of equal zero: " + num(b));
    border_ = b;
}

void FlowPropagator::setInputFlow(const Image& flow) {
    if (flow.depth() != 2)
        throw runtime_error("FlowPropagator::setInputFlow(): input flow should have depth 2: " + num(flow.depth()));
    if (flow.itemSize() != 4)
        throw runtime_error("FlowPropagator::setInputFlow(): input flow should have item size 4: " +
                            num(flow.itemSize()));
    input_ = flow;
    inputSet_ = true;
}

// FlowPropagator::compute (propagation.cu:80-134): optional negation into
// the Y buffer, then X/Y passes ping-ponging between the two buffers.
void FlowPropagator::run() {
    if (!configured_) throw logic_error("FlowPropagator::compute() stage not configured.");
    require_shape(bound_, bufX_, "FlowPropagator::compute()");
    require_alloc(bound_, "FlowPropagator::compute()");
    CrossStream x{&bound_, &bufX_, &bufY_};
    hipStream_t s = stream();
    const DevImg bx = bufX_.view(), by = bufY_.view();
    DevImg src = bound_.view();
    if (invert_) {
        launch_scale_f2(src, -1.0f, by, s);
        src = by;
    }
    launch_propagate_pass(0, src, bx, nullptr, nullptr, nullptr, nullptr, dt_, border_, s);
    launch_propagate_pass(1, bx, by, nullptr, nullptr, nullptr, nullptr, dt_, border_, s);
    for (int n = 0; n < iterations_ - 1; ++n) {
        launch_propagate_pass(0, by, bx, nullptr, nullptr, nullptr, nullptr, dt_, border_, s);
        launch_propagate_pass(1, bx, by, nullptr, nullptr, nullptr, nullptr, dt_, border_, s);
    }
    check_launch("FlowPropagator::compute");
}

void FlowPropagator::compute() {
    startTiming();
    run();
    stopTiming();
}

// ===================================================== FlowPropagatorPayload
// src/gpu/propagation.cu:217-502

FlowPropagatorPayload::FlowPropagatorPayload(const Image& flow, const Image& scalar, const Image& vector,
                                             int iterations) {
    setInputFlow(flow);
    setScalarPayload(scalar);
    setVectorPayload(vector);
    setIterations(iterations);
    configure();
}

void FlowPropagatorPayload::configure() {
    if (!flowSet_) throw runtime_error("FlowPropagatorPayload::configure(): input flow has not been set");
    if (!scalarSet_) throw runtime_error("FlowPropagatorPayload::configure(): input scalar payload has not been set");
    if (!vectorSet_) throw runtime_error("FlowPropagatorPayload::configure(): input vector payload has not been set");
    bFlow_ = flow_;
    bScalar_ = scalar_;
    bVector_ = vector_;
    const int h = bFlow_.height(), w = bFlow_.width();
    flowX_ = Image(h, w, 2, 4);
    flowY_ = Image(h, w, 2, 4);
    scalarX_ = Image(h, w, 1, 4);
    scalarY_ = Image(h, w, 1, 4);
    vectorX_ = Image(h, w, 2, 4);
    vectorY_ = Image(h, w, 2, 4);
    configured_ = true;
}

void FlowPropagatorPayload::setIterations(int n) {
    if (n <= 0) throw runtime_error("FlowPropagatorPayload::setIterations(): iterations less than zero: " + num(n));
    iterations_ = n;
    dt_ = 1.0f / float(iterations_);
}

void FlowPropagatorPayload::setBorder(int b) {
    if (b < 0)
        throw runtime_error("FlowPropagatorPayload::setBorder(): border should be greater of equal zero: " + num(b));
    border_ = b;
}

void FlowPropagatorPayload::setInputFlow(const Image& flow) {
    if (flow.depth() != 2)
        throw runtime_error("FlowPropagatorPayload::setInputFlow(): input flow should have depth 2: " +
                            num(flow.depth()));
    if (flow.itemSize() != 4)
        throw runtime_error("FlowPropagatorPayload::setInputFlow(): input flow should have item size 4: " +
                            num(flow.itemSize()));
    flow_ = flow;
    flowSet_ = true;
}

void FlowPropagatorPayload::setScalarPayload(const Image& img) {
    if (img.depth() != 1)
        throw runtime_error("FlowPropagatorPayload::setScalarPayload(): depth should be 1: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowPropagatorPayload::setScalarPayload(): item size should be 4: " +
                            num(img.itemSize()));
    if (img.height() != flow_.height() || img.width() != flow_.width())
        throw runtime_error("FlowPropagatorPayload::setScalarPayload(): scalar field shape does not match with input flow");
    scalar_ = img;
    scalarSet_ = true;
}

void FlowPropagatorPayload::setVectorPayload(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("FlowPropagatorPayload::setVectorPayload(): depth should be 2: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowPropagatorPayload::setVectorPayload(): item size should be 4: " +
                            num(img.itemSize()));
    if (img.height() != flow_.height() || img.width() != flow_.width())
        throw runtime_error("FlowPropagatorPayload::setVectorPayload(): vector field shape does not match with input flow");
    vector_ = img;
    vectorSet_ = true;
}

// FlowPropagatorPayload::compute (propagation.cu:324-373).
void FlowPropagatorPayload::run() {
    if (!configured_) throw logic_error("FlowPropagatorPayload::compute() stage not configured.");
    require_shape(bFlow_, flowX_, "FlowPropagatorPayload::compute()");
    require_shape(bScalar_, flowX_, "FlowPropagatorPayload::compute()");
    require_shape(bVector_, flowX_, "FlowPropagatorPayload::compute()");
    require_alloc(bFlow_, "FlowPropagatorPayload::compute()");
    require_alloc(bScalar_, "FlowPropagatorPayload::compute()");
    require_alloc(bVector_, "FlowPropagatorPayload::compute()");
    CrossStream x{&bFlow_, &bScalar_, &bVector_, &flowX_, &flowY_, &scalarX_, &scalarY_, &vectorX_, &vectorY_};
    hipStream_t s = stream();
    DevImg fx = flowX_.view(), fy = flowY_.view(), sx = scalarX_.view(), sy = scalarY_.view(),
           vx = vectorX_.view(), vy = vectorY_.view();
    DevImg f0 = bFlow_.view(), s0 = bScalar_.view(), v0 = bVector_.view();
    launch_propagate_pass(0, f0, fx, &s0, &sx, &v0, &vx, dt_, border_, s);
    launch_propagate_pass(1, fx, fy, &sx, &sy, &vx, &vy, dt_, border_, s);
    for (int n = 0; n < iterations_ - 1; ++n) {
        launch_propagate_pass(0, fy, fx, &sy, &sx, &vy, &vx, dt_, border_, s);
        launch_propagate_pass(1, fx, fy, &sx, &sy, &vx, &vy, dt_, border_, s);
    }
    check_launch("FlowPropagatorPayload::compute");
}

void FlowPropagatorPayload::compute() {
    startTiming();
    run();
    stopTiming();
}

// ================================================================ FlowUpdate
// src/gpu/update.cu:20-212

FlowUpdate::FlowUpdate(const Image& flow, const Image& image, const Image& gradient, float gamma, float maxflow) {
    setGamma(gamma);
    setMaxFlow(maxflow);
    setInputFlow(flow);
    setInputImage(image);
    setInputImageGradient(gradient);
    configure();
}

void FlowUpdate::configure() {
    if (!flowSet_) throw runtime_error("FlowUpdate::configure(): input flow not set");
    if (!imageSet_) throw runtime_error("FlowUpdate::configure(): input image not set");
    if (!gradSet_) throw runtime_error("FlowUpdate::configure(): input image gradient not set");
    const int h = flow_.height(), w = flow_.width();
    if (h != image_.height() || h != grad_.height() || w != image_.width() || w != grad_.width())
        throw runtime_error("FlowUpdate::configure(): input buffers do not match height and width");
    flowUpd_ = Image(h, w, 2, 4);
    imageUpd_ = Image(h, w, 1, 4);
    configured_ = true;
}

void FlowUpdate::setGamma(float g) {
    if (g <= 0) throw runtime_error("FlowUpdate::setGamma(): gamma should be greater than zero: " + std::to_string(g));
    gamma_ = g;
}

void FlowUpdate::setInputFlow(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("FlowUpdate::setInputFlow(): input flow should have depth 2: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowUpdate::setInputFlow(): input flow should have item size 4: " + num(img.itemSize()));
    flow_ = img;
    flowSet_ = true;
}

void FlowUpdate::setInputImage(const Image& img) {
    if (img.depth() != 1)
        throw runtime_error("FlowUpdate::setInputImage(): input image should have depth 1: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowUpdate::setInputImage(): input image should have item size 4: " +
                            num(img.itemSize()));
    image_ = img;
    imageSet_ = true;
}

void FlowUpdate::setInputImageGradient(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("FlowUpdate::setInputImageGradient(): input image gradient should have depth 2: " +
                            num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowUpdate::setInputImageGradient(): input image gradient should have item size 4: " +
                            num(img.itemSize()));
    grad_ = img;
    gradSet_ = true;
}

// FlowUpdate::compute (update.cu:97-116): the old image and the updated
// image are the same buffer (read then written per pixel).
void FlowUpdate::run() {
    if (!configured_) throw logic_error("FlowUpdate::compute() stage not configured.");
    require_shape(flow_, flowUpd_, "FlowUpdate::compute()");
    require_shape(image_, flowUpd_, "FlowUpdate::compute()");
    require_shape(grad_, flowUpd_, "FlowUpdate::compute()");
    require_alloc(flow_, "FlowUpdate::compute()");
    require_alloc(image_, "FlowUpdate::compute()");
    require_alloc(grad_, "FlowUpdate::compute()");
    CrossStream x{&flow_, &image_, &grad_, &imageUpd_, &flowUpd_};
    launch_update(image_.view(), grad_.view(), imageUpd_.view(), flow_.view(), imageUpd_.view(), flowUpd_.view(),
                  gamma_, maxflow_, stream());
    check_launch("FlowUpdate::compute");
}

void FlowUpdate::compute() {
    startTiming();
    run();
    stopTiming();
}

// =========================================================== DeltaFlowUpdate
// src/gpu/update.cu:221-485

DeltaFlowUpdate::DeltaFlowUpdate(const Image& flow, const Image& deltaFlow, const Image& imageOld,
                                 const Image& image, const Image& gradient, float gamma, float maxflow) {
    setGamma(gamma);
    setMaxFlow(maxflow);
    setInputDeltaFlow(deltaFlow);
    setInputFlow(flow);
    setInputImageOld(imageOld);
    setInputImage(image);
    setInputImageGradient(gradient);
    configure();
}

void DeltaFlowUpdate::configure() {
    if (!flowSet_) throw runtime_error("DeltaFlowUpdate::configure(): input flow not set");
    if (!dflowSet_) throw runtime_error("DeltaFlowUpdate::configure(): input delta flow not set");
    if (!oldSet_) throw runtime_error("DeltaFlowUpdate::configure(): input image prior not set");
    if (!imageSet_) throw runtime_error("DeltaFlowUpdate::configure(): input image not set");
    if (!gradSet_) throw runtime_error("DeltaFlowUpdate::configure(): input image gradient not set");
    const int h = dflow_.height(), w = dflow_.width();
    if (h != image_.height() || w != image_.width() || h != grad_.height() || w != grad_.width() ||
        h != imageOld_.height() || w != imageOld_.width())
        throw runtime_error("DeltaFlowUpdate::configure(): input buffers do not match height and width");
    // upper-level flow: normalized-coordinate point-sampled texture
    // (update.cu:308-313), bound here
    bFlow_ = flow_;
    const std::vector<int> xm = upper_index_map(w, bFlow_.width());
    const std::vector<int> ym = upper_index_map(h, bFlow_.height());
    xmap_.assign(xm.data(), xm.size() * sizeof(int));
    ymap_.assign(ym.data(), ym.size() * sizeof(int));
    flowUpd_ = Image(h, w, 2, 4);
    dflowUpd_ = Image(h, w, 2, 4);
    imageUpd_ = Image(h, w, 1, 4);
    configured_ = true;
}

void DeltaFlowUpdate::setGamma(float g) {
    if (g <= 0)
        throw runtime_error("DeltaFlowUpdate::setGamma(): gamma should be greater than zero: " + std::to_string(g));
    gamma_ = g;
}

void DeltaFlowUpdate::setInputFlow(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("DeltaFlowUpdate::setInputFlow(): input flow should have depth 2: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("DeltaFlowUpdate::setInputFlow(): input flow should have item size 4: " +
                            num(img.itemSize()));
    flow_ = img;
    flowSet_ = true;
}

void DeltaFlowUpdate::setInputDeltaFlow(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("DeltaFlowUpdate::setInputDeltaFlow(): input delta flow should have depth 2: " +
                            num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("DeltaFlowUpdate::setInputDeltaFlow(): input delta flow should have item size 4: " +
                            num(img.itemSize()));
    dflow_ = img;
    dflowSet_ = true;
}

void DeltaFlowUpdate::setInputImageOld(const Image& img) {
    if (img.depth() != 1)
        throw runtime_error("DeltaFlowUpdate::setInputImageOld(): input image should have depth 1: " +
                            num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("DeltaFlowUpdate::setInputImageOld(): input image should have item size 4: " +
                            num(img.itemSize()));
    imageOld_ = img;
    oldSet_ = true;
}

void DeltaFlowUpdate::setInputImage(const Image& img) {
    if (img.depth() != 1)
        throw runtime_error("DeltaFlowUpdate::setInputImage(): input image should have depth 1: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("DeltaFlowUpdate::setInputImage(): input image should have item size 4: " +
                            num(img.itemSize()));
    image_ = img;
    imageSet_ = true;
}

void DeltaFlowUpdate::setInputImageGradient(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("DeltaFlowUpdate::setInputImageGradient(): input image gradient should have depth 2: " +
                            num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error(
            "DeltaFlowUpdate::setInputImageGradient(): input image gradient should have item size 4: " +
            num(img.itemSize()));
    grad_ = img;
    gradSet_ = true;
}

// DeltaFlowUpdate::compute (update.cu:328-349).
void DeltaFlowUpdate::run() {
    if (!configured_) throw logic_error("DeltaFlowUpdate::compute() stage not configured.");
    require_shape(dflow_, flowUpd_, "DeltaFlowUpdate::compute()");
    require_shape(imageOld_, flowUpd_, "DeltaFlowUpdate::compute()");
    require_shape(image_, flowUpd_, "DeltaFlowUpdate::compute()");
    require_shape(grad_, flowUpd_, "DeltaFlowUpdate::compute()");
    for (const Image* im : {&dflow_, &imageOld_, &image_, &grad_, &bFlow_}) require_alloc(*im, "DeltaFlowUpdate::compute()");
    if (flowUpd_.height() > 0 && flowUpd_.width() > 0 && (bFlow_.height() <= 0 || bFlow_.width() <= 0))
        throw logic_error("DeltaFlowUpdate::compute(): empty upper-level flow");
    CrossStream x{&dflow_, &imageOld_, &image_, &grad_, &bFlow_, &imageUpd_, &dflowUpd_, &flowUpd_};
    launch_delta_update(image_.view(), grad_.view(), imageOld_.view(), dflow_.view(), bFlow_.view(), xmap(), ymap(),
                        imageUpd_.view(), dflowUpd_.view(), flowUpd_.view(), gamma_, maxflow_, stream());
    check_launch("DeltaFlowUpdate::compute");
}

void DeltaFlowUpdate::compute() {
    startTiming();
    run();
    stopTiming();
}

// ============================================================== FlowSmoother
// src/gpu/flowsmoothing.cu:20-152

FlowSmoother::FlowSmoother(const Image& flow, int iterations) {
    setInputFlow(flow);
    setIterations(iterations);
    configure();
}

void FlowSmoother::configure() {
    if (!inputSet_) throw runtime_error("FlowSmoother::configure(): input flow has not been set");
    bound_ = input_;
    bufX_ = Image(bound_.height(), bound_.width(), 2, 4);
    bufY_ = Image(bound_.height(), bound_.width(), 2, 4);
    configured_ = true;
}

void FlowSmoother::setIterations(int n) {
    if (n <= 0) throw runtime_error("FlowSmoother::setIterations(): iterations should be greater than zero: " + num(n));
    iterations_ = n;
}

void FlowSmoother::setInputFlow(const Image& img) {
    if (img.depth() != 2)
        throw runtime_error("FlowSmoother::setInputFlow(): input flow should have depth 2: " + num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("FlowSmoother::setInputFlow(): input flow should have item size 4: " +
                            num(img.itemSize()));
    input_ = img;
    inputSet_ = true;
}

// FlowSmoother::compute (flowsmoothing.cu:74-104).
void FlowSmoother::run() {
    if (!configured_) throw logic_error("FlowSmoother::compute() stage not configured.");
    require_shape(bound_, bufX_, "FlowSmoother::compute()");
    require_alloc(bound_, "FlowSmoother::compute()");
    CrossStream x{&bound_, &bufX_, &bufY_};
    hipStream_t s = stream();
    const DevImg bx = bufX_.view(), by = bufY_.view();
    launch_smooth_pass(0, bound_.view(), bx, s);
    launch_smooth_pass(1, bx, by, s);
    for (int n = 0; n < iterations_ - 1; ++n) {
        launch_smooth_pass(0, by, bx, s);
        launch_smooth_pass(1, bx, by, s);
    }
    check_launch("FlowSmoother::compute");
}

void FlowSmoother::compute() {
    startTiming();
    run();
    stopTiming();
}

// ============================================================== ImagePyramid
// src/gpu/pyramid.cu:22-224

ImagePyramid::ImagePyramid(const Image& image, int levels) {
    setLevels(levels);
    setInputImage(image);
    configure();
}

void ImagePyramid::configure() {
    if (!inputSet_) throw runtime_error("ImagePyramid::configure(): input image has not been set");
    bound_ = input_;
    int h = bound_.height(), w = bound_.width();
    px_.clear();
    py_.clear();
    for (int l = 0; l < levels_ - 1; ++l) {
        w /= 2;
        px_.emplace_back(h, w, 1, bound_.itemSize());
        h /= 2;
        py_.emplace_back(h, w, 1, bound_.itemSize());
    }
    configured_ = true;
}

void ImagePyramid::setInputImage(const Image& img) {
    if (img.depth() != 1) throw runtime_error("ImagePyramid::setInputImage(): image depth should be 1: " + num(img.depth()));
    if (img.itemSize() != 1 && img.itemSize() != 4)
        throw runtime_error("ImagePyramid::setInputImage(): item size should be 1 or 4: " + num(img.itemSize()));
    input_ = img;
    inputSet_ = true;
}

void ImagePyramid::setLevels(int levels) {
    if (levels <= 0) throw runtime_error("ImagePyramid::setLevels(): levels should be greater than zero: " + num(levels));
    levels_ = levels;
}

Image ImagePyramid::getImage(int level) const {
    if (level < 0 || level >= levels_)
        throw runtime_error("ImagePyramid::getImage(): level index out of range: " + num(level));
    if (level == 0) return input_;
    if ((size_t)level - 1 >= py_.size()) throw logic_error("ImagePyramid::getImage(): stage not configured");
    return py_[level - 1];
}

// ImagePyramid::compute (pyramid.cu:110-167): X then Y per level.
void ImagePyramid::run() {
    if (!configured_) throw runtime_error("ImagePyramid::compute(): stage not configured");
    require_alloc(bound_, "ImagePyramid::compute()");
    const bool u8 = bound_.itemSize() == 1;
    CrossStream x{&bound_};
    hipStream_t s = stream();
    for (size_t l = 0; l < py_.size(); ++l) {
        const Image& src = l == 0 ? bound_ : py_[l - 1];
        launch_down_pass(0, u8, src.view(), px_[l].view(), s);
        launch_down_pass(1, u8, px_[l].view(), py_[l].view(), s);
    }
    check_launch("ImagePyramid::compute");
}

void ImagePyramid::compute() {
    startTiming();
    run();
    stopTiming();
}

// =============================================================== FlowToColor
// src/gpu/display.cu:20-140

ff_image_t colorWheelRGBA() {
    size_t bytes = 0;
    const unsigned char* p = colorwheel_rgba(&bytes);
    if (bytes != (size_t)kWheelHeight * kWheelWidth * kWheelDepth)
        throw runtime_error("colour wheel table has " + num((long)bytes) + " bytes");
    ff_image_t w{};
    w.height = kWheelHeight;
    w.width = kWheelWidth;
    w.depth = kWheelDepth;
    w.itemSize = 1;
    w.pitch = (size_t)kWheelWidth * kWheelDepth;
    w.data = const_cast<unsigned char*>(p);
    return w;
}

// FlowToColor(inputFlow, maxflow) (display.cu:31-42).
FlowToColor::FlowToColor(const Image& inputFlow, float maxflow) {
    setInputFlow(inputFlow);
    setMaxFlow(maxflow);
    configure();
}

// FlowToColor::configure (display.cu:50-78): the wheel uploaded as a
// 256 x 256 RGBA image and the output colour image of the input's size.
void FlowToColor::configure() {
    if (!inputSet_) throw runtime_error("FlowToColor::configure(): input flow not set");
    const ff_image_t w = colorWheelRGBA();
    wheel_ = Image(w.height, w.width, w.depth, 1);
    wheel_.upload(w);
    color_ = Image(input_.height(), input_.width(), 4, 1);
    configured_ = true;
}

void FlowToColor::run() {
    if (!configured_) throw logic_error("FlowToColor::compute(): stage not configured");
    require_shape(input_, color_, "FlowToColor::compute()");
    require_alloc(input_, "FlowToColor::compute()");
    CrossStream x{&input_, &wheel_, &color_};
    launch_flow_to_color(input_.view(), wheel_.view(), maxflow_, color_.view(), stream());
    check_launch("FlowToColor::compute");
}

// FlowToColor::compute (display.cu:81-97).
void FlowToColor::compute() {
    startTiming();
    run();
    stopTiming();
}

// FlowToColor::setInputFlow (display.cu:100-117).
void FlowToColor::setInputFlow(const Image& flow) {
    if (flow.depth() != 2)
        throw runtime_error("FlowToColor::setInputFlow(): input flow should have depth 2: " + num(flow.depth()));
    if (flow.itemSize() != 4)
        throw runtime_error("FlowToColor::setInputFlow(): input flow should have item size 4: " + num(flow.itemSize()));
    input_ = flow;
    inputSet_ = true;
}

// FlowToColor::setMaxFlow (display.cu:130-137).
void FlowToColor::setMaxFlow(float m) {
    if (!(m > 0.0f)) throw runtime_error("FlowToColor::setMaxFlow(): maxflow should be greater than 0.0: " + std::to_string(m));
    maxflow_ = m;
}

// ===================================================== LaxWendroffPropagator
// src/gpu/propagation.cu:505-724

LaxWendroffPropagator::LaxWendroffPropagator(const Image& inputFlow, const Image& inputImage) {
    setInputFlow(inputFlow);
    setInputImage(inputImage);
    configure();
}

// LaxWendroffPropagator::configure (propagation.cu:533-569).
void LaxWendroffPropagator::configure() {
    if (!flowSet_) throw runtime_error("LaxWendroffPropagator::configure(): input flow has not been set");
    if (!imageSet_) throw runtime_error("LaxWendroffPropagator::configure(): input image has not been set");
    bufX_ = Image(flow_.height(), flow_.width(), image_.depth(), 4);
    bufY_ = Image(flow_.height(), flow_.width(), image_.depth(), 4);
    configured_ = true;
}

// LaxWendroffPropagator::compute (propagation.cu:572-644): a Y pass then an
// X pass, `iterations` times (at least once).
void LaxWendroffPropagator::run() {
    if (!configured_) throw runtime_error("LaxWendroffPropagator::compute(): stage not configured");
    require_shape(flow_, bufX_, "LaxWendroffPropagator::compute()");
    require_shape(image_, bufX_, "LaxWendroffPropagator::compute()");
    require_alloc(flow_, "LaxWendroffPropagator::compute()");
    require_alloc(image_, "LaxWendroffPropagator::compute()");
    if (image_.depth() != bufX_.depth())
        throw logic_error("LaxWendroffPropagator::compute(): input image depth changed since configure()");
    CrossStream x{&flow_, &image_, &bufX_, &bufY_};
    hipStream_t s = stream();
    const int c = image_.depth();
    launch_lax_wendroff_pass(1, c, flow_.view(), image_.view(), bufY_.view(), dt_, s);
    launch_lax_wendroff_pass(0, c, flow_.view(), bufY_.view(), bufX_.view(), dt_, s);
    for (int k = 0; k < iterations_ - 1; ++k) {
        launch_lax_wendroff_pass(1, c, flow_.view(), bufX_.view(), bufY_.view(), dt_, s);
        launch_lax_wendroff_pass(0, c, flow_.view(), bufY_.view(), bufX_.view(), dt_, s);
    }
    check_launch("LaxWendroffPropagator::compute");
}

void LaxWendroffPropagator::compute() {
    startTiming();
    run();
    stopTiming();
}

// LaxWendroffPropagator::setIterations (propagation.cu:647-658).
void LaxWendroffPropagator::setIterations(int n) {
    if (n <= 0) throw runtime_error("LaxWendroffPropagator::setIterations(): iterations less than zero: " + num(n));
    iterations_ = n;
    dt_ = 1.0f / float(iterations_);
}

// LaxWendroffPropagator::setInputFlow (propagation.cu:671-687).
void LaxWendroffPropagator::setInputFlow(const Image& flow) {
    if (flow.depth() != 2)
        throw runtime_error("LaxWendroffPropagator::setInputFlow(): input flow should have depth 2: " +
                            num(flow.depth()));
    if (flow.itemSize() != 4)
        throw runtime_error("LaxWendroffPropagator::setInputFlow(): input flow should have item size 4: " +
                            num(flow.itemSize()));
    flow_ = flow;
    flowSet_ = true;
}

// LaxWendroffPropagator::setInputImage (propagation.cu:690-714): depth 1 or
// 4, float, the input flow's size.
void LaxWendroffPropagator::setInputImage(const Image& img) {
    if (img.depth() != 1 && img.depth() != 4)
        throw runtime_error("LaxWendroffPropagator::setInputImage(): input image should have depth 1 or 4, got: " +
                            num(img.depth()));
    if (img.itemSize() != 4)
        throw runtime_error("LaxWendroffPropagator::setInputImage(): input image should have item size 4: " +
                            num(img.itemSize()));
    if (img.height() != flow_.height() || img.width() != flow_.width())
        throw runtime_error("LaxWendroffPropagator::setInputImage(): image shape " + img.shapeString() +
                            " does not match the input flow " + flow_.shapeString());
    image_ = img;
    imageSet_ = true;
}

// ============================================== RotationalFlowImagePredictor
// src/gpu/rotation.cu:20-134

RotationalFlowImagePredictor::RotationalFlowImagePredictor(const Camera& cam, const Image& inputImage) {
    setCamera(cam);
    setInputImage(inputImage);
    configure();
}

// RotationalFlowImagePredictor::configure (rotation.cu:50-68). As in the
// reference, the propagator is made anew: its iteration count returns to
// 0 (one pass pair with dt = 0) until setIterations.
void RotationalFlowImagePredictor::configure() {
    if (!imageSet_) throw runtime_error("RotationalFlowImagePredictor::configure(): input image not set");
    flow_ = Image(image_.height(), image_.width(), 2, 4);
    propagator_ = LaxWendroffPropagator(flow_, image_);
    configured_ = true;
}

// RotationalFlowImagePredictor::compute (rotation.cu:71-88).
void RotationalFlowImagePredictor::compute() {
    startTiming();
    if (!configured_) throw runtime_error("RotationalFlowImagePredictor::compute(): stage not configured");
    CrossStream x{&flow_};
    launch_rotational_flow(camera_, w_[0], w_[1], w_[2], flow_.view(), stream());
    check_launch("RotationalFlowImagePredictor::compute");
    propagator_.run();
    stopTiming();
}

// RotationalFlowImagePredictor::setInputImage (rotation.cu:91-97): the
// reference does not check the image here; configure() does (through the
// propagator).
void RotationalFlowImagePredictor::setInputImage(const Image& img) {
    image_ = img;
    imageSet_ = true;
}

// RotationalFlowImagePredictor::setAngularVelocity (rotation.cu:116-121).
void RotationalFlowImagePredictor::setAngularVelocity(float wx, float wy, float wz) {
    w_[0] = wx;
    w_[1] = wy;
    w_[2] = wz;
}

// createPerspectiveCamera (camera.cu:14-29): pixel ranges width-1 and
// height-1, focal lengths in pixels, the principal point at the centre.
Camera createPerspectiveCamera(float focalLength, int height, int width, float sensorHeight, float sensorWidth) {
    const int imgX = width - 1, imgY = height - 1;
    Camera c;
    c.alphaX = focalLength * (float)imgX / sensorWidth;
    c.alphaY = focalLength * (float)imgY / sensorHeight;
    c.centerX = 0.5f * (float)imgX;
    c.centerY = 0.5f * (float)imgY;
    return c;
}

}  // namespace ffamd
