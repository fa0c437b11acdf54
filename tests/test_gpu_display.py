"""FlowToColor (display_k.cu:14-43, display.cu:20-142) on the GPU against the
oracle's colour coding, bitwise, through the C ABI."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _flows(h, w, m, seed):
    r = np.random.default_rng(seed)
    f = r.uniform(-1.3 * m, 1.3 * m, (h, w, 2)).astype(np.float32)
    if h * w > 8:
        idx = r.integers(0, h * w, 8)
        flat = f.reshape(-1, 2)
        flat[idx[0]] = (np.nan, 0.0)
        flat[idx[1]] = (0.0, np.nan)
        flat[idx[2]] = (np.inf, -np.inf)
        flat[idx[3]] = (m, m)
        flat[idx[4]] = (-m, -m)
        flat[idx[5]] = (0.0, 0.0)
        flat[idx[6]] = (3e38, -3e38)
        flat[idx[7]] = (-0.0, 1e-30)
    return f


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (37, 131), (480, 640), (1088, 2048)])
@pytest.mark.parametrize("maxflow", [1.0, 4.0, 2.5])
def test_flow_to_color_matches_oracle(ffo, shape, maxflow):
    from flowfilter.gpu.display import FlowToColor, colorWheel
    from flowfilter.gpu.image import GPUImage
    h, w = shape
    say("flow to color", shape, maxflow)
    f = _flows(h, w, maxflow, seed=h * 7 + w)
    g = GPUImage((h, w, 2), 4)
    g.upload(f)
    c = FlowToColor(g, maxflow)
    assert c.maxflow == np.float32(maxflow)
    c.compute()
    got = c.download()
    assert got.shape == (h, w, 4) and got.dtype == np.uint8
    want = ffo.flow_to_color(f, maxflow, colorWheel())
    assert np.array_equal(got, want), np.argwhere(np.any(got != want, axis=-1))[:5]
    assert c.elapsedTime() >= 0.0


def test_flow_to_color_of_a_filter(ffo):
    """The demos' use: colour the pyramidal filter's flow every frame
    (flowWebCam.cpp:100), through the filter's device flow handle."""
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import synthetic
    from flowfilter.gpu.display import FlowToColor, colorWheel
    h, w = 272, 512
    f = gf.PyramidalFlowFilter(h, w, 2)
    f.gamma = [10.0, 50.0]
    f.maxflow = 4.0
    f.smoothIterations = [2, 2]
    o = ffo.PyramidalFlowFilter(h, w, 2)
    o.configure_like([10.0, 50.0], 4.0, [2, 2])
    color = FlowToColor(f.getFlowDevice(), 4.0)
    wheel = colorWheel()
    for t, img in enumerate(synthetic.blob_affine(h, w, 5, blobs=40, seed=21)):
        f.loadImage(img)
        f.compute()
        o.compute(img)
        color.compute()
        assert np.array_equal(color.download(), ffo.flow_to_color(o.downloadFlow(), 4.0, wheel)), t


def test_flow_to_color_errors():
    from flowfilter import errors
    from flowfilter.gpu.display import FlowToColor
    from flowfilter.gpu.image import GPUImage
    c = FlowToColor()
    with pytest.raises(errors.FilterError):
        c.configure()  # no input flow
    with pytest.raises(errors.FilterError):
        c.setInputFlow(GPUImage((4, 4), 4))  # depth 1
    with pytest.raises(errors.FilterError):
        c.maxflow = 0.0
    c.setInputFlow(GPUImage((4, 4, 2), 4))
    with pytest.raises(errors.LogicError):
        c.compute()  # not configured
    c.configure()
    c.compute()
    assert c.download().shape == (4, 4, 4)
