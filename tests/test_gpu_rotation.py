"""RotationalFlowImagePredictor and LaxWendroffPropagator (rotation.cu:20-134,
propagation.cu:505-724) on the GPU against the oracle, bitwise, through the
C ABI."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gimg(a, item=4):
    from flowfilter.gpu.image import GPUImage
    g = GPUImage(a.shape, item)
    g.upload(a)
    return g


@pytest.mark.parametrize("shape", [(1, 1), (7, 9), (120, 161), (480, 640)])
@pytest.mark.parametrize("channels", [1, 4])
@pytest.mark.parametrize("iterations", [0, 1, 3])
def test_lax_wendroff_matches_oracle(ffo, shape, channels, iterations):
    from flowfilter.gpu.propagation import LaxWendroffPropagator
    h, w = shape
    say("lax-wendroff", shape, channels, iterations)
    rng = np.random.default_rng(h * 31 + w + channels)
    flow = rng.uniform(-3, 3, (h, w, 2)).astype(np.float32)
    img = rng.uniform(0, 1, (h, w) if channels == 1 else (h, w, 4)).astype(np.float32)
    p = LaxWendroffPropagator(_gimg(flow), _gimg(img))
    if iterations:
        p.iterations = iterations
        assert p.dt == np.float32(1.0) / np.float32(iterations)
    else:
        assert p.iterations == 0 and p.dt == 0.0
    p.compute()
    got = p.getPropagatedImage().download(np.float32)
    want = ffo.lax_wendroff(flow, img, iterations)
    assert np.array_equal(got, want), float(np.abs(got - want).max())


@pytest.mark.parametrize("shape", [(9, 13), (481, 641), (1088, 2048)])
def test_rotational_predictor_matches_oracle(ffo, shape):
    from flowfilter.gpu.camera import createPerspectiveCamera
    from flowfilter.gpu.rotation import RotationalFlowImagePredictor
    h, w = shape
    say("rotational predictor", shape)
    cam = createPerspectiveCamera(4.0, h, w, 3.6, 4.8)
    camv = np.array([cam.alphaX, cam.alphaY, cam.centerX, cam.centerY], np.float32)
    assert np.array_equal(camv, ffo.create_perspective_camera(4.0, h, w, 3.6, 4.8))
    img = np.random.default_rng(7).uniform(0, 1, (h, w)).astype(np.float32)
    p = RotationalFlowImagePredictor(cam, _gimg(img))
    assert p.iterations == 0
    for (wx, wy, wz), it in [((0.0, 0.0, 0.0), 2), ((0.01, -0.02, 0.005), 4), ((-0.003, 0.001, 0.03), 1)]:
        p.setAngularVelocity(wx, wy, wz)
        p.iterations = it
        p.compute()
        flow = p.getOpticalFlow().download(np.float32)
        want_flow = ffo.rotational_flow(camv, wx, wy, wz, h, w)
        assert np.array_equal(flow, want_flow), (wx, wy, wz)
        pred = p.getPredictedImage().download(np.float32)
        assert np.array_equal(pred, ffo.lax_wendroff(want_flow, img, it)), (wx, wy, wz, it)
    assert p.elapsedTime() >= 0.0
    # the camera the predictor uses can be read back and replaced
    cam2 = createPerspectiveCamera(6.0, h, w, 3.6, 4.8)
    p.setCamera(cam2)
    assert p.getCamera().alphaX == cam2.alphaX
    # configure() makes the propagator anew: iterations back to 0
    p.configure()
    assert p.iterations == 0


def test_rotation_and_lax_wendroff_errors():
    from flowfilter import errors
    from flowfilter.gpu.camera import createPerspectiveCamera
    from flowfilter.gpu.image import GPUImage
    from flowfilter.gpu.propagation import LaxWendroffPropagator
    from flowfilter.gpu.rotation import RotationalFlowImagePredictor
    p = LaxWendroffPropagator()
    with pytest.raises(errors.FilterError):
        p.configure()
    with pytest.raises(errors.FilterError):
        p.setInputImage(GPUImage((4, 4), 4))  # flow not set: shapes differ
    p.setInputFlow(GPUImage((4, 4, 2), 4))
    with pytest.raises(errors.FilterError):
        p.setInputImage(GPUImage((4, 4, 2), 4))  # depth 2
    with pytest.raises(errors.FilterError):
        p.setInputImage(GPUImage((4, 4), 1))  # u8
    with pytest.raises(errors.FilterError):
        p.iterations = 0
    with pytest.raises(errors.FilterError):
        p.compute()  # not configured
    r = RotationalFlowImagePredictor(createPerspectiveCamera(4.0, 8, 8, 3.6, 4.8))
    with pytest.raises(errors.FilterError):
        r.compute()  # not configured
    r.setInputImage(GPUImage((8, 8), 1))
    with pytest.raises(errors.FilterError):
        r.configure()  # u8 image: the propagator refuses it
    r.setInputImage(GPUImage((8, 8), 4))
    r.configure()
    r.compute()
