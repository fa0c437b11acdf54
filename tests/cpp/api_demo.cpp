// api_demo.cpp -- C++ use of libflowfilter_amd through include/flowfilter_amd.hpp,
// written the way a caller of the reference's flowfilter::gpu classes would
// (the timing loop of demos/flowProfiler): build a PyramidalFlowFilter, feed
// frames, read the flow back, check the exception mapping.
//
// Usage: api_demo [height width levels frames]
// Prints one line "ok <mean ms> <max |flow|>" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "flowfilter_amd.hpp"

int main(int argc, char** argv) {
    using namespace flowfilter;
    const int h = argc > 1 ? std::atoi(argv[1]) : 240;
    const int w = argc > 2 ? std::atoi(argv[2]) : 320;
    const int levels = argc > 3 ? std::atoi(argv[3]) : 2;
    const int frames = argc > 4 ? std::atoi(argv[4]) : 10;

    gpu::PyramidalFlowFilter filter(h, w, levels);
    std::vector<float> gamma(levels, 10.0f);
    filter.setGamma(gamma);
    filter.setMaxFlow(4.0f);
    std::vector<int> smooth(levels, 2);
    filter.setSmoothIterations(smooth);

    image_t img = createImage(h, w, 1, 1);
    image_t flow = createImage(h, w, 2, sizeof(float));
    double total = 0.0;
    for (int t = 0; t < frames; ++t) {
        unsigned char* px = static_cast<unsigned char*>(img.data);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const double u = x - 0.8 * t, v = y + 0.3 * t;
                px[y * img.pitch + x] = (unsigned char)(127.5 + 60 * std::sin(0.21 * u) * std::cos(0.17 * v));
            }
        filter.loadImage(img);
        filter.compute();
        total += filter.elapsedTime();
    }
    filter.downloadFlow(flow);
    float mx = 0.0f;
    const float* f = static_cast<const float*>(flow.data);
    for (int i = 0; i < h * w * 2; ++i) mx = std::fmax(mx, std::fabs(f[i]));

    // shape mismatch -> std::invalid_argument (src/gpu/image.cu:117)
    bool threw = false;
    image_t bad = createImage(h, w + 1, 1, 1);
    try {
        filter.loadImage(bad);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    destroyImage(bad);
    // level out of range -> std::runtime_error (std::exception in the reference)
    bool threw2 = false;
    try {
        filter.setGamma(levels, 1.0f);
    } catch (const std::runtime_error&) {
        threw2 = true;
    }
    destroyImage(img);
    destroyImage(flow);
    if (!threw || !threw2 || !(mx <= 4.0f)) {
        std::printf("FAILED threw=%d threw2=%d max=%g\n", threw, threw2, mx);
        return 1;
    }
    std::printf("ok %.4f %.4f\n", total / frames, mx);
    return 0;
}
