"""flowfilter.gpu -- GPU pipeline stages and filters (HIP, gfx950).

Module layout follows the reference's python/flowfilter/gpu package:
image, imagemodel, propagation, update, pyramid, flowfilters; plus
flowsmoothing (FlowSmoother, which the reference exposes only in C++).
"""
