"""Pins for the oracle's area-sampled card coverage and its adjoint (CPU).

The coverage of a pixel is the exact area of triangle ∩ pixel box, so the
forward is pinned by closed forms (axis-aligned squares, fractional
offsets, a triangle covering the image), by the total being the triangle's
area, by brute-force supersampling and by projective invariance.  The
backward is pinned by the closed-form gradient of a triangle's area and by
central finite differences of the float64 forward; a wrong normal sign,
swapped edge weight or missing w-term fails one of them.
"""
import numpy as np
import pytest

import oracle


def clip_of(points_px, H, W, w=1.0):
    p = np.asarray(points_px, np.float64)
    out = np.zeros((1, p.shape[0], 4), np.float64)
    out[0, :, 0] = (2.0 * p[:, 0] / W - 1.0) * w
    out[0, :, 1] = (1.0 - 2.0 * p[:, 1] / H) * w
    out[0, :, 3] = w
    return out


def random_tris(rng, n, H, W, margin=2.0):
    pts = rng.uniform([margin, margin], [W - margin, H - margin], size=(n * 3, 2))
    return clip_of(pts, H, W), np.arange(n * 3, dtype=np.int32).reshape(n, 3)


def tri_area_px(pos, tri, H, W):
    x = 0.5 * (1 + pos[0, :, 0] / pos[0, :, 3]) * W
    y = 0.5 * (1 - pos[0, :, 1] / pos[0, :, 3]) * H
    a, b, c = tri.T
    return 0.5 * np.abs((x[b] - x[a]) * (y[c] - y[a]) - (y[b] - y[a]) * (x[c] - x[a]))


QUAD = np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def test_aligned_square_is_integer():
    H = W = 8
    pos = clip_of([(2, 1), (5, 1), (5, 4), (2, 4)], H, W)
    c = oracle.card_coverage((H, W), pos, QUAD)[0]
    want = np.zeros((H, W))
    want[1:4, 2:5] = 1.0
    np.testing.assert_array_equal(c, want)


def test_fractional_offsets():
    H = W = 8
    pos = clip_of([(2.5, 1.25), (5.5, 1.25), (5.5, 4.25), (2.5, 4.25)], H, W)
    c = oracle.card_coverage((H, W), pos, QUAD)[0]
    col = np.array([0.5, 1.0, 1.0, 0.5])       # x-overlap of columns 2..5
    row = np.array([0.75, 1.0, 1.0, 0.25])     # y-overlap of rows 1..4
    want = np.zeros((H, W))
    want[1:5, 2:6] = row[:, None] * col[None, :]
    np.testing.assert_allclose(c, want, atol=1e-12)


def test_orientation_and_vertex_order_do_not_matter():
    H = W = 16
    rng = np.random.default_rng(3)
    pos, tri = random_tris(rng, 6, H, W)
    a = oracle.card_coverage((H, W), pos, tri)
    b = oracle.card_coverage((H, W), pos, tri[:, ::-1].copy())
    c = oracle.card_coverage((H, W), pos, np.roll(tri, 1, axis=1).copy())
    np.testing.assert_allclose(a, b, atol=1e-12)
    np.testing.assert_allclose(a, c, atol=1e-12)


def test_total_is_triangle_area():
    H = W = 32
    rng = np.random.default_rng(0)
    pos, tri = random_tris(rng, 20, H, W)
    for t in range(tri.shape[0]):
        c = oracle.card_coverage((H, W), pos, tri[t:t + 1])
        np.testing.assert_allclose(c.sum(), tri_area_px(pos, tri[t:t + 1], H, W)[0], rtol=1e-12)


def test_triangle_covering_image_gives_ones():
    H, W = 6, 10
    pos = clip_of([(-50, -50), (200, -50), (-50, 200)], H, W)
    c = oracle.card_coverage((H, W), pos, np.array([[0, 1, 2]], np.int32))
    np.testing.assert_allclose(c, 1.0, atol=1e-12)


def test_matches_supersampling():
    H = W = 12
    rng = np.random.default_rng(1)
    pos, tri = random_tris(rng, 3, H, W, margin=1.0)
    c = oracle.card_coverage((H, W), pos, tri)[0]
    n = 64
    s = (np.arange(n) + 0.5) / n
    x = 0.5 * (1 + pos[0, :, 0]) * W
    y = 0.5 * (1 - pos[0, :, 1]) * H
    gx = (np.arange(W)[:, None] + s[None]).reshape(-1)
    gy = (np.arange(H)[:, None] + s[None]).reshape(-1)
    X, Y = np.meshgrid(gx, gy)
    inside_sum = np.zeros_like(X)
    for a, b, cc in tri:
        def side(i, j):
            return (x[j] - x[i]) * (Y - y[i]) - (y[j] - y[i]) * (X - x[i])
        s0, s1, s2 = side(a, b), side(b, cc), side(cc, a)
        inside_sum += ((s0 >= 0) & (s1 >= 0) & (s2 >= 0)) | ((s0 <= 0) & (s1 <= 0) & (s2 <= 0))
    brute = inside_sum.reshape(H, n, W, n).mean(axis=(1, 3))
    assert np.abs(brute - c).max() < 3.0 / n
    assert abs(brute.sum() - c.sum()) < 0.5 * c.sum() / n


def test_projective_scale_invariance_and_w_guard():
    H = W = 16
    rng = np.random.default_rng(2)
    pos, tri = random_tris(rng, 5, H, W)
    a = oracle.card_coverage((H, W), pos, tri)
    b = oracle.card_coverage((H, W), pos * 2.5, tri)
    np.testing.assert_allclose(a, b, atol=1e-12)
    neg = pos.copy()
    neg[0, tri[0, 0]] *= -1.0  # one vertex behind the eye: triangle skipped
    d = oracle.card_coverage((H, W), neg, tri)
    e = oracle.card_coverage((H, W), pos, tri[1:])
    np.testing.assert_allclose(d, e, atol=1e-12)


def test_window_and_selection():
    H = W = 24
    rng = np.random.default_rng(4)
    pos, tri = random_tris(rng, 8, H, W)
    full = oracle.card_coverage((H, W), pos, tri, sel=[1, 5])
    sub = oracle.card_coverage((H, W), pos, tri[[1, 5]])
    np.testing.assert_allclose(full, sub, atol=1e-12)
    win = oracle.card_coverage((H, W), pos, tri, sel=[1, 5],
                               window=(np.array([3]), np.array([7]), 10, 12))
    np.testing.assert_allclose(win[0], full[0, 3:13, 7:19], atol=1e-12)


def test_grad_of_total_area_closed_form():
    """d(sum C)/d(vertex) for a triangle inside the image is the gradient of
    its area: 1/2 the opposite edge rotated outward, chained to clip space."""
    H, W = 20, 30
    rng = np.random.default_rng(5)
    for _ in range(5):
        pos, tri = random_tris(rng, 1, H, W)
        w = rng.uniform(0.5, 2.0, size=3)
        pos[0, :, :2] *= w[:, None]
        pos[0, :, 3] = w
        g = oracle.card_coverage_grad((H, W), pos, tri, np.ones((1, H, W)))
        x = 0.5 * (1 + pos[0, :, 0] / pos[0, :, 3]) * W
        y = 0.5 * (1 - pos[0, :, 1] / pos[0, :, 3]) * H
        cr = (x[1] - x[0]) * (y[2] - y[0]) - (y[1] - y[0]) * (x[2] - x[0])
        sgn = np.sign(cr)
        for i in range(3):
            j, k = (i + 1) % 3, (i + 2) % 3
            gx = 0.5 * sgn * (y[j] - y[k])
            gy = 0.5 * sgn * (x[k] - x[j])
            P = pos[0, i]
            want = np.array([gx * 0.5 * W / P[3], -gy * 0.5 * H / P[3], 0.0,
                             (-gx * 0.5 * W * P[0] + gy * 0.5 * H * P[1]) / P[3] ** 2])
            np.testing.assert_allclose(g[0, i], want, rtol=1e-9, atol=1e-9)


def _loss(res, pos, tri, target):
    return ((oracle.card_coverage(res, pos, tri) - target) ** 2).sum()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grad_matches_central_differences(seed):
    H, W = 24, 20
    rng = np.random.default_rng(seed)
    pos, tri = random_tris(rng, 6, H, W)
    # shared vertices: a strip of triangles over 8 points plus w != 1
    pts = rng.uniform([3, 3], [W - 3, H - 3], size=(8, 2))
    pos = clip_of(pts, H, W)
    w = rng.uniform(0.7, 1.5, size=8)
    pos[0, :, :2] *= w[:, None]
    pos[0, :, 3] = w
    tri = np.array([[0, 1, 2], [1, 3, 2], [2, 3, 4], [3, 5, 4], [4, 5, 6], [5, 7, 6]], np.int32)
    target = rng.uniform(0, 1.5, size=(1, H, W))
    cov = oracle.card_coverage((H, W), pos, tri)
    g = oracle.card_coverage_grad((H, W), pos, tri, 2 * (cov - target))
    h = 1e-6
    fd = np.zeros_like(g)
    for v in range(8):
        for k in (0, 1, 3):
            d = np.zeros_like(pos)
            d[0, v, k] = h
            fd[0, v, k] = (_loss((H, W), pos + d, tri, target) -
                           _loss((H, W), pos - d, tri, target)) / (2 * h)
    err = np.linalg.norm(fd - g) / np.linalg.norm(fd)
    assert err < 1e-5, err
    assert np.all(g[..., 2] == 0)


def test_loss_difference_helper_matches_full_recompute():
    H = W = 32
    rng = np.random.default_rng(6)
    pos, tri = random_tris(rng, 10, H, W)
    target = rng.uniform(0, 1, size=(1, H, W))
    cov = oracle.card_coverage((H, W), pos, tri)
    d = np.zeros_like(pos)
    d[0, tri[3, 1]] = [1e-3, -2e-3, 0, 5e-4]
    got = oracle.coverage_loss_difference((H, W), pos, tri, target, cov, d)
    want = 0.5 * (_loss((H, W), pos + d, tri, target) - _loss((H, W), pos - d, tri, target))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
