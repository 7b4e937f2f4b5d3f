"""Pins for the MSAA texture-patch oracle and the sampling host logic (CPU).

The filter and blur weights are pinned by closed forms (Mitchell-Netravali
at 0, 1 and 2 and its partition of unity; the Gaussian taps), the blur by
constant, impulse and border images, and the patch by a fully covered
patch, an empty one, and a half-plane whose far pixels are fully in or out.
"""
import math

import numpy as np
import pytest

import oracle
from strns2cards_amd import texture


def test_mitchell_netravali_closed_forms():
    f = texture.mitchell_netravali
    assert f(0.0) == pytest.approx(8 / 9)
    assert f(1.0) == pytest.approx(1 / 18)
    assert f(-1.0) == pytest.approx(1 / 18)
    assert f(2.0) == 0.0 and f(2.5) == 0.0
    for x in np.linspace(0, 1, 11):  # partition of unity over integer shifts
        assert sum(f(x + k) for k in range(-3, 4)) == pytest.approx(1.0, abs=1e-12)


def test_sample_grid_and_weights():
    offs, w = texture.msaa_samples(8, 2.0)
    assert len(offs) == 64 and offs[0] == (-0.875, -0.875) and offs[1] == (-0.625, -0.875)
    assert sum(w) == pytest.approx(1.0, abs=1e-12)
    W = np.array(w).reshape(8, 8)
    np.testing.assert_allclose(W, W.T, atol=1e-15)
    np.testing.assert_allclose(W, W[::-1, ::-1], atol=1e-15)


def test_gaussian_taps():
    k = texture.gaussian_kernel3(1.5)
    e = math.exp(-1 / 4.5)
    np.testing.assert_allclose(k, [e / (1 + 2 * e), 1 / (1 + 2 * e), e / (1 + 2 * e)], rtol=1e-6)


def test_blur_constant_impulse_border():
    k = texture.gaussian_kernel3(1.5)
    c = np.full((5, 7), 0.75, np.float32)
    np.testing.assert_allclose(oracle.blur3x3_reflect(c, k), c, rtol=1e-6)
    imp = np.zeros((5, 7), np.float32)
    imp[2, 3] = 1.0
    out = oracle.blur3x3_reflect(imp, k)
    np.testing.assert_allclose(out[1:4, 2:5], np.outer(k, k), rtol=1e-6)
    assert out.sum() == pytest.approx(1.0, rel=1e-6)
    corner = np.zeros((4, 4), np.float32)
    corner[0, 0] = 1.0  # reflect: the corner's neighbours see it once, itself k1^2
    out = oracle.blur3x3_reflect(corner, k)
    assert out[0, 0] == pytest.approx(k[1] * k[1], rel=1e-6)
    assert out[0, 1] == pytest.approx(k[0] * k[1], rel=1e-6)
    assert out[1, 1] == pytest.approx(k[0] * k[0], rel=1e-6)


def _setup(H, W):
    offs, w = texture.msaa_samples()
    dx = np.array([o[0] * 2.0 / W for o in offs], np.float32)
    dy = np.array([o[1] * 2.0 / H for o in offs], np.float32)
    return dx, dy, np.array(w, np.float32), texture.gaussian_kernel3(1.5)


def test_patch_full_and_empty():
    H, W = 12, 16
    dx, dy, w, k = _setup(H, W)
    pos = np.array([[-3, -3, 0], [3, -3, 0], [3, 3, 0], [-3, 3, 0]], np.float32)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    attr = np.tile(np.array([0.2, -0.4, 1.0], np.float32), (4, 1))
    rgb, alpha = oracle.msaa_patch(pos, tri, attr, H, W, dx, dy, w, k)
    np.testing.assert_allclose(rgb, np.broadcast_to([0.6, 0.3, 1.0], rgb.shape), atol=1e-6)
    np.testing.assert_allclose(alpha, 1.0, atol=1e-5)
    rgb, alpha = oracle.msaa_patch(pos, np.zeros((0, 3), np.int32), attr, H, W, dx, dy, w, k)
    np.testing.assert_array_equal(rgb, np.broadcast_to(np.float32([0.5, 1.0, 0.5]), rgb.shape))
    np.testing.assert_array_equal(alpha, 0.0)


def test_patch_half_plane():
    H, W = 8, 32
    dx, dy, w, k = _setup(H, W)
    # left half of NDC covered: pixels more than 2 px left of x = 0 are fully
    # covered by all samples, more than 2 px right of it by none
    pos = np.array([[-3, -3, 0], [0, -3, 0], [0, 3, 0], [-3, 3, 0]], np.float32)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    attr = np.tile(np.array([1.0, 0.0, -1.0], np.float32), (4, 1))
    rgb, alpha = oracle.msaa_patch(pos, tri, attr, H, W, dx, dy, w, k)
    np.testing.assert_allclose(alpha[:, :12], 1.0, atol=1e-5)
    np.testing.assert_array_equal(alpha[:, 20:], 0.0)
    np.testing.assert_allclose(rgb[:, :12], np.broadcast_to([1.0, 0.5, 0.0], (H, 12, 3)), atol=1e-6)
    assert np.all(np.diff(alpha[3]) <= 1e-6)  # monotone across the edge


def test_patch_mask_segment():
    # a horizontal segment at NDC y = 0.5 across x in [-0.5, 0.5] of an
    # 8 x 16 patch: pixel row 0.5 (1 - 0.5) 8 = 2, flipped to row 5,
    # columns 4 .. 11 drawn at (1 - t1) + t1 (255, or 254 when it rounds
    # below 1), nothing else
    pts = np.array([[-0.5, 0.5, 0.0], [0.5, 0.5, 0.0]], np.float32)
    m = oracle.patch_mask(pts, 1, 2, 8, 16)
    assert m.dtype == np.uint8 and m.shape == (8, 16)
    assert np.all(m[5, 4:12] >= 254)
    m[5, 4:12] = 0
    assert m[5, 12] == 0 or m[5, 12] >= 254  # the end pixel x = 12.0 itself
    m[5, 12] = 0
    assert not m.any()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_packing_frame_equals_per_wisp_reductions(dt):
    """The segmented min / max gives the reference's per-wisp extents
    (texture_utils.py:128-141) exactly, in the same dtypes."""
    rng = np.random.default_rng(0)
    wisps = [np.cumsum(rng.normal(size=(int(rng.integers(1, 6)), int(rng.integers(1, 40)), 3)),
                       1).astype(dt) for _ in range(40)]
    _, _, pb, sc = texture.packing_frame(wisps)
    pts = np.concatenate([w.reshape(-1, 3) for w in wisps])
    pb_ref = pts[:, :2].max(axis=0) - pts[:, :2].min(axis=0)
    size = np.zeros(3)
    for w in wisps:
        p = w.reshape(-1, 3)
        size = np.maximum(size, p.max(axis=0) - p.min(axis=0))
    sc_ref = 2.0 * np.ones(3) / size
    sc_ref[2] *= 0.2
    assert pb.dtype == pb_ref.dtype and sc.dtype == sc_ref.dtype
    np.testing.assert_array_equal(pb, pb_ref)
    np.testing.assert_array_equal(sc, sc_ref)
