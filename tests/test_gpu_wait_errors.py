"""A fused-frame wait that gives up is an error, never a silent FF_OK.

The level below a pyramid's top waits inside the fused launch for the top
level's tiles; that wait is bounded by a spin limit. Here the test hook
(ff_filter_set_wait_test) makes every such wait target a count no launch
reaches, with a spin limit of a few polls, so each wait gives up at once and
every kernel still ends normally. The next synchronizing call (elapsedTime,
getFlow/getImage, which download) must then raise FilterError
(FF_ERR_RUNTIME); later calls without new timeouts succeed again.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H, W = 272, 512  # both levels run in the fused frame kernel


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pyramid():
    import flowfilter.gpu.flowfilters as gf
    f = gf.PyramidalFlowFilter(H, W, 2)
    f.gamma = [10.0, 50.0]
    f.maxflow = 4.0
    f.smoothIterations = [2, 2]
    return f


def _frames(n, seed=5):
    from flowfilter import synthetic
    return list(synthetic.blob_affine(H, W, n, blobs=40, seed=seed))


def test_timed_out_wait_raises_on_next_sync():
    from flowfilter import errors
    say("wait test: single pyramid")
    f = _pyramid()
    frames = _frames(6)
    for img in frames[:3]:  # first frame staged, then fused frames
        f.loadImage(img)
        f.compute()
    f.getFlow()
    f.elapsedTime()
    assert f.waitTimeouts() == 0
    f.setWaitTest(4)
    f.loadImage(frames[3])
    f.compute()
    with pytest.raises(errors.FilterError, match="spin limit"):
        f.elapsedTime()
    assert f.waitTimeouts() > 0
    # reported once: the next call without new timeouts succeeds
    f.elapsedTime()
    f.getFlow()
    # a download reports it too
    f.loadImage(frames[4])
    f.compute()
    with pytest.raises(errors.FilterError, match="spin limit"):
        f.getFlow()
    f.getImage()
    # the normal wait again: no error
    f.setWaitTest(0)
    n = f.waitTimeouts()
    f.loadImage(frames[5])
    f.compute()
    f.getFlow()
    f.elapsedTime()
    assert f.waitTimeouts() == n


def test_timed_out_wait_in_a_batch_raises_for_batch_and_members():
    import flowfilter.gpu.flowfilters as gf
    from flowfilter import errors
    say("wait test: batch")
    a, b = _pyramid(), _pyramid()
    batch = gf.BatchedFlowFilter([a, b])
    frames = _frames(4, seed=6)
    for img in frames[:2]:
        a.loadImage(img)
        b.loadImage(img)
        batch.compute()
    batch.elapsedTime()
    a.getFlow()
    b.getFlow()
    batch.setWaitTest(4)
    a.loadImage(frames[2])
    b.loadImage(frames[2])
    batch.compute()
    with pytest.raises(errors.FilterError, match="spin limit"):
        batch.elapsedTime()
    with pytest.raises(errors.FilterError, match="spin limit"):
        a.getFlow()
    with pytest.raises(errors.FilterError, match="spin limit"):
        b.getImage()
    batch.setWaitTest(0)
    a.loadImage(frames[3])
    b.loadImage(frames[3])
    batch.compute()
    batch.elapsedTime()
    assert np.array_equal(a.getFlow(), b.getFlow())
