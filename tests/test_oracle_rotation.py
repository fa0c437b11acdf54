"""Pins the oracle's rotational flow (rotation_k.cu:17-43, camera.cu:14-29)
and Lax-Wendroff propagation (propagation_k.h:49-132, propagation.cu:572-644)
by what the mathematics fixes:

  * camera: focal lengths in pixels and the principal point at the centre;
  * rotational flow: zero angular velocity gives zero flow; at the principal
    point the flow is (ax wy, -ay wx); pure roll, yaw and pitch give the
    textbook rotational fields (Longuet-Higgins & Prazdny), which a sign or
    index mistake in the cross product would break;
  * Lax-Wendroff: constants are invariant for any flow; linear and quadratic
    images under a constant flow are translated exactly (the scheme is exact
    for polynomials of degree <= 2; the test values are dyadic, so float
    arithmetic is exact too); the Y pass runs before the X pass (an image
    constant along y under a flow that varies along y tells the orders
    apart); no iterations set means one pass pair with dt = 0 (the input).
"""
import numpy as np
import pytest

F32 = np.float32


def test_camera(ffo):
    c = ffo.create_perspective_camera(4.0, 480, 640, 3.6, 4.8)
    assert c[2] == F32(319.5) and c[3] == F32(239.5)
    assert abs(c[0] - 4.0 * 639 / 4.8) < 1e-3 and abs(c[1] - 4.0 * 479 / 3.6) < 1e-3


def _grid(h, w, cam):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    return (x - cam[2]) / cam[0], (y - cam[3]) / cam[1], x, y


@pytest.mark.parametrize("shape", [(33, 41), (120, 160)])
def test_rotational_flow_closed_forms(ffo, shape):
    h, w = shape
    cam = ffo.create_perspective_camera(3.0, h, w, 2.0, 2.5).astype(np.float64)
    ax, ay, cx, cy = cam
    px, py, x, y = _grid(h, w, cam)
    assert np.array_equal(ffo.rotational_flow(cam, 0.0, 0.0, 0.0, h, w), np.zeros((h, w, 2), F32))
    # principal point (an integer pixel for odd sizes): flow (ax wy, -ay wx)
    f = ffo.rotational_flow(cam, 0.25, -0.5, 0.125, h, w)
    if h % 2 and w % 2:
        assert np.array_equal(f[h // 2, w // 2], np.array([ax * -0.5, -ay * 0.25], F32))
    tol = 2e-5

    def close(got, u, v):
        ref = np.stack([u, v], -1)
        return np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())

    wz = 0.02  # roll: u = -wz ax py, v = wz ay px
    assert close(ffo.rotational_flow(cam, 0.0, 0.0, wz, h, w), -wz * ax * py, wz * ay * px)
    wy = 0.01  # yaw: u = wy ax (1 + px^2), v = wy ay px py
    assert close(ffo.rotational_flow(cam, 0.0, wy, 0.0, h, w), wy * ax * (1 + px * px), wy * ay * px * py)
    wx = 0.01  # pitch: u = -wx ax px py, v = -wx ay (1 + py^2)
    assert close(ffo.rotational_flow(cam, wx, 0.0, 0.0, h, w), -wx * ax * px * py, -wx * ay * (1 + py * py))


def _const_flow(h, w, u, v):
    f = np.zeros((h, w, 2), F32)
    f[..., 0], f[..., 1] = u, v
    return f


@pytest.mark.parametrize("channels", [1, 4])
def test_lax_wendroff_invariants(ffo, channels):
    h, w = 24, 30
    rng = np.random.default_rng(1)
    flow = rng.uniform(-2, 2, (h, w, 2)).astype(F32)
    shape = (h, w) if channels == 1 else (h, w, 4)
    const = np.full(shape, 0.7, F32)
    for it in (0, 1, 3):
        assert np.array_equal(ffo.lax_wendroff(flow, const, it), const)
    img = rng.uniform(0, 1, shape).astype(F32)
    assert np.array_equal(ffo.lax_wendroff(flow, img, 0), img)  # dt = 0


@pytest.mark.parametrize("channels", [1, 4])
@pytest.mark.parametrize("iterations", [1, 2, 4])
def test_lax_wendroff_translates_polynomials_exactly(ffo, channels, iterations):
    h, w = 20, 26
    u, v = 0.75, -0.5
    flow = _const_flow(h, w, u, v)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    m = iterations + 2  # interior: untouched by the clamped edges
    cases = [(3 * x - 2 * y + 5, 3 * (x - u) - 2 * (y - v) + 5),
             (x * x + 0.5 * y * y, (x - u) ** 2 + 0.5 * (y - v) ** 2)]
    for img64, want64 in cases:
        img = img64.astype(F32)
        if channels == 4:
            img = np.repeat(img[..., None], 4, -1) * np.array([1, 2, 0.5, -1], F32)
            want64 = want64[..., None] * np.array([1, 2, 0.5, -1])
        got = ffo.lax_wendroff(flow, img, iterations)
        assert np.array_equal(got[m:-m, m:-m], want64[m:-m, m:-m].astype(F32))


def test_lax_wendroff_pass_order(ffo):
    """Y pass first: an image constant along y (I = x) is left unchanged by
    the Y pass, so one pair with dt = 1 gives x - u(x, y) inside, exactly,
    even where u varies along y; X first would then move the varying
    result along y."""
    h, w = 16, 20
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    flow = np.zeros((h, w, 2), F32)
    flow[..., 0] = (0.25 * (y % 4)).astype(F32)
    flow[..., 1] = 0.5
    img = x.astype(F32)
    got = ffo.lax_wendroff(flow, img, 1)
    assert np.array_equal(got[2:-2, 2:-2], (x - flow[..., 0])[2:-2, 2:-2].astype(F32))
    # the other order differs
    xy = ffo.lax_wendroff_pass(1, flow, ffo.lax_wendroff_pass(0, flow, img, 1.0), 1.0)
    assert not np.array_equal(xy[2:-2, 2:-2], got[2:-2, 2:-2])


def test_lax_wendroff_axes_are_transposes(ffo):
    """The X pass on an image equals the transpose of the Y pass on the
    transposed image with the flow components swapped (exact for images
    whose differences are exact in float)."""
    rng = np.random.default_rng(4)
    img = rng.integers(0, 64, (12, 17)).astype(F32)
    flow = (rng.integers(-8, 9, (12, 17, 2)) / 8.0).astype(F32)
    fx = ffo.lax_wendroff_pass(0, flow, img, 0.5)
    ft = np.ascontiguousarray(flow.transpose(1, 0, 2)[..., ::-1])
    fy = ffo.lax_wendroff_pass(1, ft, np.ascontiguousarray(img.T), 0.5)
    assert np.array_equal(fx, fy.T)
