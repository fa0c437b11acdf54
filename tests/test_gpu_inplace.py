"""Frames read in place (PyramidalFlowFilter.setImageDevice): the fused
frame kernel loads the caller's device frame directly instead of a copy
into the filter's input image.

Bar: flow and brightness bitwise equal to the same filter fed by
loadImageDevice (a copy), for 1 to 4 levels (4: copied in, deeper pyramids
read the pyramid images), for aligned frames and for a frame address that
is not a multiple of 4 (copied in); no copy launch when read in place."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(h, w, levels):
    import flowfilter.gpu.flowfilters as gf
    f = gf.PyramidalFlowFilter(h, w, levels)
    f.gamma = [10.0, 30.0, 50.0, 70.0][:levels] if levels > 1 else 10.0
    f.maxflow = 4.0
    f.smoothIterations = [2, 2, 1, 1][:levels] if levels > 1 else 2
    return f


@pytest.mark.parametrize("levels", [1, 2, 3, 4])
@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("shape", [(272, 480), (301, 411)])
def test_frames_in_place_match_copies(levels, offset, shape):
    from flowfilter import _native as N, synthetic
    from flowfilter.gpu.image import GPUImage
    (H, W), n = shape, 6
    frames = list(synthetic.translating_noise(H, W, n, (1.1, -0.7), seed=40 + levels))
    stack = np.concatenate(frames, axis=0)
    g = GPUImage((n * H, W + 4), 1)
    g.upload(np.ascontiguousarray(np.pad(stack, ((0, 0), (offset, 4 - offset)))))
    a, b = make(H, W, levels), make(H, W, levels)
    for t in range(n):
        say("in place: levels", levels, "offset", offset, "frame", t)
        ptr = g.data + t * H * g.pitch + offset
        N.profile_begin("copy")
        a.setImageDevice(ptr, g.pitch)
        a.compute()
        _, copies = N.profile_end()
        b.loadImageDevice(ptr, g.pitch)
        b.compute()
        if t >= 1 and levels <= 3 and offset == 0:
            assert copies == 0, (t, copies)
        assert np.array_equal(a.getFlow(), b.getFlow()), (levels, offset, t)
        assert np.array_equal(a.getImage(), b.getImage()), (levels, offset, t)


def test_later_load_replaces_frame_in_place():
    from flowfilter import synthetic
    from flowfilter.gpu.image import GPUImage
    H, W = 128, 192
    frames = list(synthetic.translating_noise(H, W, 4, (0.8, 0.5), seed=7))
    junk = GPUImage((H, W), 1)
    junk.upload(np.full((H, W), 17, np.uint8))
    a, b = make(H, W, 2), make(H, W, 2)
    for t in range(4):
        a.setImageDevice(junk.data, junk.pitch)
        a.loadImage(frames[t])  # replaces the frame given in place
        a.compute()
        b.loadImage(frames[t])
        b.compute()
        assert np.array_equal(a.getFlow(), b.getFlow()), t


def test_batch_copies_frames_given_in_place():
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import synthetic
    from flowfilter.gpu.image import GPUImage
    H, W = 128, 192
    frames = list(synthetic.translating_noise(H, W, 4, (0.8, 0.5), seed=8))
    g = GPUImage((4 * H, W), 1)
    g.upload(np.concatenate(frames, axis=0))
    a, b = make(H, W, 2), make(H, W, 2)
    batch = gf.BatchedFlowFilter([a])
    for t in range(4):
        a.setImageDevice(g.data + t * H * g.pitch, g.pitch)
        batch.compute()
        b.loadImage(frames[t])
        b.compute()
        assert np.array_equal(a.getFlow(), b.getFlow()), t


def test_set_image_device_errors():
    from flowfilter import errors
    f = make(64, 64, 2)
    with pytest.raises(errors.InvalidArgument):
        f.setImageDevice(0, 64)
    with pytest.raises(errors.InvalidArgument):
        f.setImageDevice(4096, 32)


def test_frame_in_place_beyond_32_bit_offsets_is_copied():
    """A frame in place whose pitch x height reaches 2^31 bytes lies beyond
    the fused kernel's 32-bit buffer offsets: compute() copies it into the
    input image instead (one copy launch per frame), with the flow of the
    same frames loaded from the host."""
    import torch
    from flowfilter import _native as N, synthetic
    H, W = 272, 480
    pitch = ((2 ** 31 // H) + 256) // 256 * 256
    assert pitch * H >= 2 ** 31
    say("frame in place with pitch", pitch)
    frames = list(synthetic.translating_noise(H, W, 3, (1.1, -0.7), seed=9))
    big = torch.zeros((H, pitch), dtype=torch.uint8, device="cuda")
    a, b = make(H, W, 2), make(H, W, 2)
    for t in range(3):
        big[:, :W] = torch.from_numpy(frames[t]).cuda()
        torch.cuda.synchronize()
        N.profile_begin("copy")
        a.setImageDevice(big.data_ptr(), pitch)
        a.compute()
        _, copies = N.profile_end()
        b.loadImage(frames[t])
        b.compute()
        if t >= 1:
            assert copies == 1, (t, copies)
        assert np.array_equal(a.getFlow(), b.getFlow()), t
        assert np.array_equal(a.getImage(), b.getImage()), t
