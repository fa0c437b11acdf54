"""flowfilter -- Adarve-Mahony optical-flow filter on AMD MI355X (gfx950).

Drop-in for the ``flowfilter`` Python package's GPU modules
(python/flowfilter/gpu/*.pyx of the reference): the same module names,
classes, methods and properties, bound to the C ABI of libflowfilter_amd.so
instead of the CUDA library.

    import flowfilter.gpu.flowfilters as gpufilter
    f = gpufilter.PyramidalFlowFilter(480, 640, 2)
    f.gamma = [10, 50]; f.maxflow = 4; f.smoothIterations = [2, 2]
    f.loadImage(img); f.compute(); flow = f.getFlow()
"""
__version__ = "1.0.0"
