"""Pins for the CPU oracle (runs without a GPU).

The reference ships no tests, fixtures or golden vectors, so each oracle
function is checked against what the mathematics fixes: closed forms on
hand-built inputs, brute force on tiny inputs, and invariants chosen so a
plausible restatement mistake (dropped term, wrong sign, swapped vertex,
off-by-one clamp) fails at least one check.
"""
import numpy as np
import pytest

import oracle

# ------------------------------------------------------------- helpers


def ndc_x(px, W):
    return 2.0 * px / W - 1.0


def ndc_y(py, H):
    return 1.0 - 2.0 * py / H


def clip(points_px, H, W, z=0.0, w=1.0):
    """Pixel-space (x, y) points -> clip space with depth z and weight w."""
    p = np.asarray(points_px, np.float64)
    out = np.zeros((1, p.shape[0], 4), np.float32)
    out[0, :, 0] = ndc_x(p[:, 0], W) * w
    out[0, :, 1] = ndc_y(p[:, 1], H) * w
    out[0, :, 2] = np.broadcast_to(z, p.shape[0]) * w
    out[0, :, 3] = w
    return out


def seg_frags(p0, p1, H=64, W=64, z=0.0):
    pos = clip([p0, p1], H, W, z)
    return oracle.dda_compute_fragments((H, W), pos, np.array([[0, 1]], np.int32))


# ------------------------------------------------------------------ DDA


def test_dda_horizontal_closed_form():
    fp, fa = seg_frags((3.25, 5.5), (10.75, 5.5))
    np.testing.assert_array_equal(fp[:, 2], np.arange(3, 11))
    assert (fp[:, 1] == 5).all() and (fp[:, 0] == 0).all()
    assert (fa[:, 0] == 0).all()  # x-major
    assert (fa[:, 3] == 1).all()  # edge id + 1
    centres = np.arange(3, 11) + 0.5
    np.testing.assert_allclose(fa[:, 1], (centres - 3.25) / 7.5, atol=1e-6)


def test_dda_reversed_endpoints_same_pixels_t_measured_from_v0():
    fp, fa = seg_frags((10.75, 5.5), (3.25, 5.5))
    np.testing.assert_array_equal(fp[:, 2], np.arange(3, 11))
    centres = np.arange(3, 11) + 0.5
    np.testing.assert_allclose(fa[:, 1], (10.75 - centres) / 7.5, atol=1e-6)


def test_dda_vertical_and_diagonal():
    fp, fa = seg_frags((7.5, 2.2), (7.5, 9.9))
    np.testing.assert_array_equal(fp[:, 1], np.arange(2, 10))
    assert (fp[:, 2] == 7).all() and (fa[:, 0] == 1).all()
    fp, _ = seg_frags((2.5, 2.5), (9.5, 9.5))
    np.testing.assert_array_equal(fp[:, 1], fp[:, 2])
    np.testing.assert_array_equal(fp[:, 2], np.arange(2, 10))


def test_dda_zero_length_offscreen_and_clipping():
    assert seg_frags((4.3, 4.3), (4.3, 4.3))[0].shape[0] == 0
    assert seg_frags((-30.0, 5.5), (-2.0, 5.5))[0].shape[0] == 0
    assert seg_frags((3.0, 70.5), (40.0, 70.5))[0].shape[0] == 0
    fp, _ = seg_frags((-10.2, 5.5), (70.7, 5.5))
    np.testing.assert_array_equal(fp[:, 2], np.arange(0, 64))
    # counts from the span pass: floor(x1) - floor(x0) + 1, clipped
    _, _, _, n = oracle.dda_compute_span((64, 64), clip([(-10.2, 5.5), (70.7, 5.5)], 64, 64),
                                         np.array([[0, 1]], np.int32))
    assert n == 64


def test_dda_depth_clip():
    assert seg_frags((3.5, 5.5), (20.5, 5.5), z=1.5)[0].shape[0] == 0
    pos = clip([(0.5, 5.5), (40.5, 5.5)], 64, 64)
    pos[0, 0, 2], pos[0, 1, 2] = -0.5, 1.5  # z crosses 1 at t = 0.75
    fp, fa = oracle.dda_compute_fragments((64, 64), pos, np.array([[0, 1]], np.int32))
    assert fa[:, 2].max() <= 1.0 and fa[:, 2].min() >= -1.0
    assert fp[:, 2].max() == 30  # centre 30.5 -> t = 0.75 exactly, z = 1.0 kept


def test_dda_random_invariants():
    """Fragments are consecutive along the major axis, stay within one pixel
    of the segment, and start in the first endpoint's major pixel."""
    rng = np.random.default_rng(0)
    H = W = 64
    for _ in range(300):
        a, b = rng.uniform(1, 63, 2), rng.uniform(1, 63, 2)
        fp, fa = seg_frags(a, b, H, W)
        if fp.shape[0] == 0:
            continue
        major = 0 if abs(b[0] - a[0]) >= abs(b[1] - a[1]) else 1
        m = fp[:, 2] if major == 0 else fp[:, 1]
        lo = min(a[major], b[major])
        assert np.all(np.diff(m) == 1)
        assert m[0] == int(np.floor(lo))
        minor = fp[:, 1] if major == 0 else fp[:, 2]
        assert np.all(np.abs(np.diff(minor)) <= 1)
        # distance of the pixel centre to the infinite line <= sqrt(2)/2 + 1
        d = b - a
        n = np.array([-d[1], d[0]]) / np.linalg.norm(d)
        c = np.stack([fp[:, 2] + 0.5, fp[:, 1] + 0.5], 1)
        assert np.abs((c - a) @ n).max() <= 1.0 + 1e-4


def test_dda_perspective_t1_reprojects():
    """t1 is the clip-space parameter whose projection is the sample's
    major-axis coordinate (perspective-correct interpolation)."""
    rng = np.random.default_rng(1)
    H, W = 48, 80
    for _ in range(50):
        P = np.zeros((1, 2, 4), np.float32)
        wv = rng.uniform(0.5, 2.5, 2)
        P[0, :, 0] = rng.uniform(-0.9, 0.9, 2) * wv
        P[0, :, 1] = rng.uniform(-0.9, 0.9, 2) * wv
        P[0, :, 2] = rng.uniform(-0.5, 0.5, 2) * wv
        P[0, :, 3] = wv
        fp, fa = oracle.dda_compute_fragments((H, W), P, np.array([[0, 1]], np.int32))
        for (b, h, w_), (axis, t1, z, _) in zip(fp, fa):
            q = (1 - t1) * P[0, 0].astype(np.float64) + t1 * P[0, 1]
            if axis == 0:
                xs = 2 * (q[0] / q[3] + 1) / 2 * W / 2  # pixel x of the clip point
                assert w_ <= xs + 1e-3 and xs - 1e-3 <= w_ + 1
            else:
                ys = (1 - q[1] / q[3]) / 2 * H
                assert h <= ys + 1e-3 and ys - 1e-3 <= h + 1
            assert abs(q[2] / q[3] - z) <= 1e-5


# ------------------------------------------------------------ triangles


def brute_cover(P, H, W):
    """float64 inclusive edge-function coverage of one triangle, w = 1."""
    pts = set()
    (x0, y0), (x1, y1), (x2, y2) = P
    A = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    for y in range(H):
        for x in range(W):
            px, py = x + 0.5, y + 0.5
            w0 = ((x1 - px) * (y2 - py) - (x2 - px) * (y1 - py)) / A
            w1 = ((x2 - px) * (y0 - py) - (x0 - px) * (y2 - py)) / A
            if w0 >= -1e-9 and w1 >= -1e-9 and w0 + w1 <= 1 + 1e-9:
                pts.add((y, x, min(abs(w0), abs(w1), abs(1 - w0 - w1))))
    return pts


def test_triangle_brute_force_and_barycentrics():
    rng = np.random.default_rng(2)
    H, W = 24, 32
    for _ in range(40):
        P = rng.uniform(-2, 34, size=(3, 2))
        pos = clip(P, H, W)
        fp, fa = oracle.compute_fragments((H, W), pos, np.array([[0, 1, 2]], np.int32))
        got = {(int(y), int(x)) for _, y, x in fp}
        want = brute_cover(P, H, W)
        sure = {(y, x) for y, x, m in want if m > 1e-5}
        maybe = {(y, x) for y, x, m in want}
        assert sure <= got <= maybe | {(y, x) for y, x in got if (y, x) in maybe}
        assert got <= maybe
        # barycentric reconstruction of the pixel centre (screen space)
        for (_, y, x), (b0, b1, z, tid) in zip(fp, fa):
            c = b0 * P[0] + b1 * P[1] + (1 - b0 - b1) * P[2]
            assert abs(c[0] - (x + 0.5)) < 1e-3 and abs(c[1] - (y + 0.5)) < 1e-3
            assert tid == 1


def test_triangle_perspective_barycentrics():
    """(sum b_i p_i)_xy / (sum b_i w_i) reprojects to the pixel centre."""
    rng = np.random.default_rng(3)
    H, W = 20, 28
    P = np.zeros((1, 3, 4), np.float32)
    w = rng.uniform(0.6, 2.0, 3)
    P[0, :, 0] = np.array([-0.8, 0.9, 0.1]) * w
    P[0, :, 1] = np.array([-0.7, -0.5, 0.9]) * w
    P[0, :, 2] = 0.1 * w
    P[0, :, 3] = w
    fp, fa = oracle.compute_fragments((H, W), P, np.array([[0, 1, 2]], np.int32))
    assert fp.shape[0] > 50
    Pd = P[0].astype(np.float64)
    for (_, y, x), (b0, b1, z, _) in zip(fp, fa):
        q = b0 * Pd[0] + b1 * Pd[1] + (1 - b0 - b1) * Pd[2]
        assert abs(q[0] / q[3] - (-1 + 2 * (x + 0.5) / W)) < 1e-5
        assert abs(q[1] / q[3] - (1 - 2 * (y + 0.5) / H)) < 1e-5
        assert abs(z - 0.1) < 1e-5


def test_triangle_right_angle_count_shared_edge_and_degenerate():
    H = W = 16
    # Square [2,10]^2 split on the diagonal through pixel centres: both
    # triangles contain the 8 diagonal pixels (inclusive tests).
    sq = clip([(2, 2), (10, 2), (10, 10), (2, 10)], H, W)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    fp, fa = oracle.compute_fragments((H, W), sq, tri)
    t0 = {(y, x) for (_, y, x), a in zip(fp, fa) if a[3] == 1}
    t1 = {(y, x) for (_, y, x), a in zip(fp, fa) if a[3] == 2}
    diag = {(k, k) for k in range(2, 10)}
    assert t0 & t1 == diag
    assert len(t0 | t1) == 64
    assert len(t0) == 36 and len(t1) == 36  # 28 strict + 8 shared each
    # degenerate (collinear) triangle: A == 0 -> nothing
    line = clip([(1, 1), (5, 5), (9, 9)], H, W)
    assert oracle.compute_fragments((H, W), line, np.array([[0, 1, 2]], np.int32))[0].shape[0] == 0


def test_depth_test_nearest_and_ties():
    rng = np.random.default_rng(4)
    n = 500
    fp = np.stack([rng.integers(0, 2, n), rng.integers(0, 5, n), rng.integers(0, 6, n)], 1).astype(np.int32)
    fa = rng.uniform(-1, 1, size=(n, 4)).astype(np.float32)
    fa[::7, 2] = 0.25  # exact ties
    fa[1::11, 2] = 1e-9  # z + 2 rounds to 2.0: ties with z = 0 ...
    fa[2::11, 2] = 0.0
    rast = oracle.depth_test(2, (5, 6), fp, fa)
    best = {}
    for i in range(n):
        key = (np.float32(fa[i, 2]) + np.float32(2.0), i)  # quantized depth, then index
        k = tuple(fp[i])
        if k not in best or key < best[k][0]:
            best[k] = (key, i)
    for (b, y, x), (_, i) in best.items():
        np.testing.assert_array_equal(rast[b, y, x], fa[i])
    assert (rast.reshape(-1, 4)[[i for i in range(60) if False]] == 0).all()
    covered = np.zeros((2, 5, 6), bool)
    covered[fp[:, 0], fp[:, 1], fp[:, 2]] = True
    assert (rast[~covered] == 0).all()


# -------------------------------------------------------------- bitsets


def test_cluster_mask_order_free_and_exact():
    rng = np.random.default_rng(5)
    n, K = 3000, 70
    fp = np.stack([rng.integers(0, 3, n), rng.integers(0, 9, n), rng.integers(0, 11, n)], 1).astype(np.int32)
    fa = np.zeros((n, 4), np.float32)
    fa[:, 3] = rng.integers(1, 41, n)
    p2c = rng.integers(0, K, 40).astype(np.int32)
    bits = oracle.cluster_mask_from_fragments((9, 11), fp, fa, p2c, K, B=3)
    perm = rng.permutation(n)
    np.testing.assert_array_equal(bits, oracle.cluster_mask_from_fragments((9, 11), fp[perm], fa[perm],
                                                                           p2c, K, B=3))
    want = np.zeros((3, 9, 11, 3), np.uint32)
    for (b, y, x), a in zip(fp, fa):
        k = p2c[int(a[3]) - 1]
        want[b, y, x, k // 32] |= np.uint32(1 << (k % 32))
    np.testing.assert_array_equal(bits.view(np.uint32), want)


def test_popcount_and_rgb_closed_forms():
    rng = np.random.default_rng(6)
    bits = rng.integers(0, 2**32, size=(2, 3, 4, 3), dtype=np.uint64).astype(np.uint32).view(np.int32)
    cnt = oracle.popcount_bitset(bits)
    want = np.vectorize(lambda v: bin(int(v) & 0xFFFFFFFF).count("1"))(bits).sum(-1)
    np.testing.assert_array_equal(cnt, want)
    # integer-valued colours keep every fp32 sum exact: rgb[k] = (k, 1, 2^(k%8))
    k = np.arange(96)
    rgb = np.stack([k, np.ones_like(k), 2.0 ** (k % 8)], 1).astype(np.float32)
    out = oracle.accumulate_bitset_rgb(bits, rgb)
    u = bits.view(np.uint32)
    for idx in np.ndindex(2, 3, 4):
        ks = [s * 32 + j for s in range(3) for j in range(32) if (int(u[idx][s]) >> j) & 1]
        assert out[idx][0] == sum(ks) and out[idx][1] == len(ks)
        assert out[idx][2] == sum(2.0 ** (q % 8) for q in ks)


def test_compaction_is_stable_filter():
    rng = np.random.default_rng(7)
    a = rng.integers(-3, 100, 1000).astype(np.int32)
    np.testing.assert_array_equal(oracle.compact_valid_ints(a), a[a >= 0])
    fp = np.stack([a, a + 1, a + 2], 1).astype(np.int32)
    fa = rng.normal(size=(1000, 4)).astype(np.float32)
    o_fp, o_fa = oracle.filter_valid_fragments(fp, fa)
    np.testing.assert_array_equal(o_fp, fp[a >= 0])
    np.testing.assert_array_equal(o_fa, fa[a >= 0])


# ------------------------------------------------------------ silhouette


def test_silhouette_marking():
    H = W = 32
    # single triangle: every edge is a boundary edge
    tri = np.array([[0, 1, 2]], np.int32)
    P = clip([(4, 4), (20, 6), (8, 24)], H, W)
    edges = np.array([[0, 1], [1, 2], [2, 0]], np.int32)
    e2t = np.array([[0, -1], [0, -1], [0, -1]], np.int32)
    prim, normals = oracle.mark_discontinuity_edges((H, W), P, tri, edges, e2t)
    np.testing.assert_array_equal(prim, [0, 1, 2])
    np.testing.assert_allclose(np.linalg.norm(normals, axis=1), 1, atol=1e-6)
    centroid = np.array([(4 + 20 + 8) / 3, (4 + 6 + 24) / 3])
    pts = np.array([(4, 4), (20, 6), (8, 24)], np.float64)
    for e, (a, b) in enumerate(edges):
        mid = (pts[a] + pts[b]) / 2
        assert normals[e] @ (centroid - mid) < 0  # points away from the triangle
    # flat quad: diagonal not a silhouette; folded quad: it is
    quad_tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    e = np.array([[0, 2], [0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    e2t = np.array([[0, 1], [0, -1], [0, -1], [1, -1], [1, -1]], np.int32)
    flat = clip([(4, 4), (20, 4), (20, 20), (4, 20)], H, W)
    prim, _ = oracle.mark_discontinuity_edges((H, W), flat, quad_tri, e, e2t)
    np.testing.assert_array_equal(prim, [1, 2, 3, 4])
    folded = clip([(4, 4), (20, 4), (20, 20), (14, 6)], H, W)  # v3 folded over the diagonal
    prim, _ = oracle.mark_discontinuity_edges((H, W), folded, quad_tri, e, e2t)
    assert 0 in prim


# ------------------------------------------------------------ antialias


def _rect_card(x0, y0, x1, y1, H, W):
    pos = clip([(x0, y0), (x1, y0), (x1, y1), (x0, y1)], H, W)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    edges = np.array([[0, 2], [0, 1], [1, 2], [2, 3], [0, 3]], np.int32)
    e2t = np.array([[0, 1], [0, -1], [0, -1], [1, -1], [1, -1]], np.int32)
    return pos, tri, edges, e2t


def _grad(card_rect, target_rect, rho=1.0, H=48, W=48, R=7.0):
    pos, tri, edges, e2t = _rect_card(*card_rect, H, W)
    e2c = np.zeros(edges.shape[0], np.int32)
    bits = oracle.triangle_cluster_mask((H, W), pos, tri, np.zeros(2, np.int32), 1)
    tpos, ttri, _, _ = _rect_card(*target_rect, H, W)
    tgt = oracle.triangle_cluster_mask((H, W), tpos, ttri, np.zeros(2, np.int32), 1)
    return oracle.antialiased_cluster_bitset_grad(bits, tgt, pos, tri, edges, e2t, e2c, R, rho)


def test_antialias_zero_when_matching():
    g = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 30.3, 33.3))
    assert np.abs(g).max() == 0


def test_antialias_sign_and_chain_rule():
    """Card larger than target on the right: the right vertices get
    d/dx > 0 (loss grows when the edge moves out), others ~ 0 in x."""
    g = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3))
    gx = g[0, :, 0]
    assert gx[1] > 0 and gx[2] > 0  # vertices on x = 30.3
    assert abs(gx[0]) < 1e-6 * gx[1] and abs(gx[3]) < 1e-6 * gx[1]
    assert (g[0, :, 2] == 0).all()  # no depth gradient
    # card smaller than target on the left: left vertices get d/dx > 0
    g2 = _grad((14.3, 12.3, 30.3, 33.3), (10.3, 12.3, 30.3, 33.3))
    assert g2[0, 0, 0] > 0 and g2[0, 3, 0] > 0


def test_antialias_rho_scales_false_positives_only():
    fp1 = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3), rho=1.0)
    fp5 = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3), rho=0.5)
    np.testing.assert_allclose(fp5, 0.5 * fp1, rtol=1e-6, atol=0)
    fn1 = _grad((14.3, 12.3, 30.3, 33.3), (10.3, 12.3, 30.3, 33.3), rho=1.0)
    fn5 = _grad((14.3, 12.3, 30.3, 33.3), (10.3, 12.3, 30.3, 33.3), rho=0.5)
    np.testing.assert_array_equal(fn5, fn1)


def test_antialias_translation_equivariance():
    """x / y gradients move with the geometry; the w gradient does not,
    because antialias.cu:269 weights it by the absolute sample position
    (dpw = (dx*px + dy*py) / pw^2)."""
    a = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3))
    b = _grad((13.3, 7.3, 33.3, 28.3), (13.3, 7.3, 27.3, 28.3))
    np.testing.assert_allclose(a[..., :3], b[..., :3], rtol=1e-4, atol=1e-6 * np.abs(a).max())


def test_antialias_w_gradient_uses_absolute_position():
    """For w = 1 the w-gradient of each silhouette sample is
    (dx*px + dy*py) split by (t0, t1).  Vertex 1 (30.3, 12.3) gets x from the
    vertical right edge (px constant) and y from the horizontal bottom edge
    (py constant), so g_w[1] = px_right*g_x[1] + py_bottom*g_y[1]; likewise
    vertex 2 with the top edge."""
    H = W = 48
    g = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3), H=H, W=W)
    px = 2 * 30.3 / W - 1
    py_bottom, py_top = 1 - 2 * 12.3 / H, 1 - 2 * 33.3 / H
    np.testing.assert_allclose(g[0, 1, 3], px * g[0, 1, 0] + py_bottom * g[0, 1, 1], rtol=1e-4)
    np.testing.assert_allclose(g[0, 2, 3], px * g[0, 2, 0] + py_top * g[0, 2, 1], rtol=1e-4)
    assert abs(g[0, 1, 1]) > 0 and abs(g[0, 2, 1]) > 0


def test_antialias_hand_computed_right_edge():
    """Vertical right edge at x = 30.3 from y = 12.3 to 33.3, target 6 px
    narrower, empty elsewhere: every DDA sample sits at (30.3, k + 0.5), its
    own pixel (30) lies outside the card so c01 = 1, and dc is the Gaussian
    mass (sigma = R/4, cut at radius R) of the false-positive strip
    x in [24.3, 30.3) inside the window.  d/dx of the two vertices sums to
    sum_k dc_k * W/2 (w = 1, normal (+1, 0))."""
    H = W = 48
    R = 7.0
    g = _grad((10.3, 12.3, 30.3, 33.3), (10.3, 12.3, 24.3, 33.3), H=H, W=W, R=R)
    sig = R / 4
    total = 0.0
    for k in range(12, 34):  # pixel rows whose centre lies in [12.3, 33.3]
        yc = k + 0.5
        if not (12.3 <= yc <= 33.3):
            continue
        dc = 0.0
        for y in range(max(k - 7, 0), min(k + 7, H - 1) + 1):
            for x in range(30 - 7, 30 + 8):
                cx, cy = x + 0.5, y + 0.5
                inside_card = 10.3 <= cx <= 30.3 and 12.3 <= cy <= 33.3
                inside_tgt = 10.3 <= cx <= 24.3 and 12.3 <= cy <= 33.3
                d2 = (cx - 30.3) ** 2 + (cy - yc) ** 2
                if d2 < R * R:
                    dc += np.exp(-d2 / (2 * sig * sig)) / (2 * np.pi * sig * sig) * (inside_card - inside_tgt)
        total += dc * W / 2
    np.testing.assert_allclose(g[0, 1, 0] + g[0, 2, 0], total, rtol=1e-4)


# --------------------------------------------------------------- curves


def test_smoothing_fixed_points_and_identity():
    n = 60
    s = np.linspace(0, 1, n)
    x = np.stack([s, 2 * s, -s], 1)[None].astype(np.float32)
    arc = oracle.compute_arclength(x)
    beta = 0.02
    out = oracle.smooth_strands(x, arc, beta)
    inner = (arc[0] > 3 * beta * np.sqrt(6) + 1e-3) & (arc[0] < arc[0, -1] - 3 * beta * np.sqrt(6) - 1e-3)
    np.testing.assert_allclose(out[0, inner], x[0, inner], atol=2e-6)
    np.testing.assert_array_equal(oracle.smooth_strands(x, arc, 1e-6), x)  # only j = i survives
    c = np.broadcast_to(np.array([0.3, -1.0, 2.0], np.float32), (1, n, 3)).copy()
    np.testing.assert_allclose(oracle.smooth_strands(c, arc, 0.1), c, rtol=1e-6)


def test_skinning_identity_and_rigid_rotation():
    rng = np.random.default_rng(8)
    x = rng.normal(size=(5, 30, 3)).astype(np.float32).cumsum(1)
    arc = oracle.compute_arclength(x)
    eye = np.broadcast_to(np.eye(3, dtype=np.float32), (5, 30, 3, 3)).copy()
    h = rng.normal(size=x.shape).astype(np.float32)
    np.testing.assert_allclose(oracle.skinning(x, arc, h, h, eye, 2.0), x, atol=2e-5)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Rm = np.broadcast_to(q.astype(np.float32), (5, 30, 3, 3)).copy()
    z = np.zeros_like(x)
    np.testing.assert_allclose(oracle.skinning(x, arc, z, z, Rm, 2.0), x @ q.T.astype(np.float32),
                               atol=2e-5)
    # total weight 0 (beta tiny and points far apart) keeps x_i
    np.testing.assert_array_equal(oracle.skinning(x, arc, h, h, eye, 1e-30), x)


@pytest.mark.parametrize("n", [0, 1, 2048, 5000])
def test_span_prefix_is_inclusive_cumsum(n):
    rng = np.random.default_rng(n)
    pos = clip(rng.uniform(-5, 70, size=(max(n, 1) + 1, 2)), 64, 64)
    edges = np.stack([np.arange(n), np.arange(n) + 1], 1).astype(np.int32)
    prefix, slopes, spans, total = oracle.dda_compute_span((64, 64), pos, edges)
    counts = [oracle.dda_compute_span((64, 64), pos, edges[i:i + 1])[3] for i in range(min(n, 50))]
    np.testing.assert_array_equal(prefix[:len(counts)], np.cumsum(counts))
    assert total == (prefix[-1] if n else 0)


# ------------------------------------------------- attribute interpolation
# dda.cu:258-342 and rasterize.cu:367-472: image = sum of the primitive's
# vertex attributes weighted by (t0, t1) or (b0, b1, 1 - b0 - b1), with the
# primitive decoded as int(rast[..., 3]) - 1; the backward scatters the same
# weights.  Hand-built rasters with asymmetric attributes fix the weight
# order, the vertex order and the id decode; rasters from the DDA and the
# triangle raster fix that the weights reconstruct the sample point; float64
# adjoint identities fix the backwards against the forwards.


def _asym_attr(V, C, seed):
    """Attributes with no two vertices or channels alike."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-3.0, 3.0, size=(V, C)).astype(np.float32)
    a += np.arange(V, dtype=np.float32)[:, None] * 10.0
    a += np.arange(C, dtype=np.float32)[None, :] * 0.125
    return a


def test_dda_interpolate_hand_built_raster():
    edges = np.array([[0, 1], [3, 5], [4, 2]], np.int32)
    attr = _asym_attr(6, 3, 10)
    rast = np.zeros((1, 2, 3, 4), np.float32)
    rast[0, 0, 0] = [0, 0.25, 0.1, 2.0]  # edge 1 = (3, 5), t1 = 0.25
    rast[0, 1, 2] = [1, 0.75, 0.3, 3.0]  # edge 2 = (4, 2), t1 = 0.75
    rast[0, 0, 1] = [0, 0.5, 0.0, 1.0]   # edge 0 = (0, 1), t1 = 0.5
    img = oracle.dda_interpolate_attributes(rast, attr, edges).astype(np.float64)
    a = attr.astype(np.float64)
    np.testing.assert_allclose(img[0, 0, 0], 0.75 * a[3] + 0.25 * a[5], rtol=1e-6)
    np.testing.assert_allclose(img[0, 1, 2], 0.25 * a[4] + 0.75 * a[2], rtol=1e-6)
    np.testing.assert_allclose(img[0, 0, 1], 0.5 * a[0] + 0.5 * a[1], rtol=1e-6)
    empty = np.ones((1, 2, 3), bool)
    empty[0, 0, 0] = empty[0, 1, 2] = empty[0, 0, 1] = False
    assert (img[empty] == 0).all()  # rast[..., 3] == 0: nothing written


def test_dda_interpolate_reconstructs_sample_point():
    """Interpolating the screen coordinates of the vertices over a DDA raster
    gives the fragment's sample point: on the pixel centre's major-axis
    coordinate for samples strictly inside the segment, on the segment, and
    inside the pixel."""
    rng = np.random.default_rng(11)
    H, W = 40, 48
    P = rng.uniform(1.0, 47.0, size=(12, 2))
    P[:, 1] = np.clip(P[:, 1], 1.0, 39.0)
    pos = clip(P, H, W, z=rng.uniform(-0.5, 0.5, size=12))
    edges = np.array([[2 * i, 2 * i + 1] for i in range(6)], np.int32)
    fp, fa = oracle.dda_compute_fragments((H, W), pos, edges)
    rast = oracle.depth_test(1, (H, W), fp, fa)
    img = oracle.dda_interpolate_attributes(rast, P.astype(np.float32), edges).astype(np.float64)
    hit = np.argwhere(rast[0, :, :, 3] > 0)
    assert hit.shape[0] > 60
    interior = 0
    for y, x in hit:
        e = int(rast[0, y, x, 3]) - 1
        axis = int(rast[0, y, x, 0])
        q = img[0, y, x]
        p0, p1 = P[edges[e, 0]], P[edges[e, 1]]
        # on the segment (collinear, between the endpoints)
        d = p1 - p0
        cross = d[0] * (q[1] - p0[1]) - d[1] * (q[0] - p0[0])
        assert abs(cross) <= 1e-3 * np.hypot(*d)
        s = np.dot(q - p0, d) / np.dot(d, d)
        assert -1e-5 <= s <= 1 + 1e-5
        # inside the fragment's pixel (closed square)
        assert x - 1e-3 <= q[0] <= x + 1 + 1e-3 and y - 1e-3 <= q[1] <= y + 1 + 1e-3
        c = (x + 0.5) if axis == 0 else (y + 0.5)
        lo, hi = sorted((p0[axis], p1[axis]))
        if lo < c < hi:
            interior += 1
            assert abs(q[axis] - c) < 1e-4
    assert interior > 40


def test_dda_interpolate_backward_hand_built_and_adjoint():
    edges = np.array([[0, 1], [3, 5], [4, 2]], np.int32)
    rast = np.zeros((1, 2, 2, 4), np.float32)
    rast[0, 0, 0] = [0, 0.25, 0.0, 2.0]
    rast[0, 1, 1] = [0, 0.5, 0.0, 2.0]  # the same edge again: contributions add
    g = np.zeros((1, 2, 2, 2), np.float32)
    g[0, 0, 0] = [1.0, -2.0]
    g[0, 1, 1] = [4.0, 8.0]
    d, ab = oracle.backward_dda_interpolate_attributes(rast, g, edges, 6, with_abs=True)
    want = np.zeros((6, 2))
    want[3] = 0.75 * np.array([1.0, -2.0]) + 0.5 * np.array([4.0, 8.0])
    want[5] = 0.25 * np.array([1.0, -2.0]) + 0.5 * np.array([4.0, 8.0])
    np.testing.assert_allclose(d.astype(np.float64), want, rtol=1e-6, atol=0)
    # the sum of |term| each element carries (the scale of its rounding)
    want_abs = np.zeros((6, 2))
    want_abs[3] = 0.75 * np.array([1.0, 2.0]) + 0.5 * np.array([4.0, 8.0])
    want_abs[5] = 0.25 * np.array([1.0, 2.0]) + 0.5 * np.array([4.0, 8.0])
    np.testing.assert_allclose(ab.astype(np.float64), want_abs, rtol=1e-6, atol=0)
    # <interp(a), g> == <a, interp_bwd(g)> on a raster from the DDA
    rng = np.random.default_rng(12)
    H, W = 32, 32
    P = rng.uniform(1.0, 31.0, size=(16, 2))
    pos = clip(P, H, W)
    edges = np.array([[i, i + 1] for i in range(15)], np.int32)
    fp, fa = oracle.dda_compute_fragments((H, W), pos, edges)
    rast = oracle.depth_test(1, (H, W), fp, fa)
    a = rng.normal(size=(16, 3)).astype(np.float32)
    g = rng.normal(size=(1, H, W, 3)).astype(np.float32)
    lhs = (oracle.dda_interpolate_attributes(rast, a, edges).astype(np.float64) * g).sum()
    rhs = (a.astype(np.float64) *
           oracle.backward_dda_interpolate_attributes(rast, g, edges, 16)).sum()
    scale = np.abs(g).sum() * np.abs(a).max()
    assert abs(lhs - rhs) <= 1e-6 * scale


def test_triangle_interpolate_hand_built_raster():
    tri = np.array([[0, 1, 2], [5, 3, 4]], np.int32)
    attr = _asym_attr(6, 4, 13)
    rast = np.zeros((2, 2, 2, 4), np.float32)
    rast[0, 1, 0] = [0.5, 0.25, 0.3, 2.0]    # triangle 1 = (5, 3, 4)
    rast[1, 0, 1] = [0.125, 0.625, 0.1, 1.0]  # triangle 0 = (0, 1, 2)
    img = oracle.interpolate_triangle_attributes(rast, attr, tri).astype(np.float64)
    a = attr.astype(np.float64)
    np.testing.assert_allclose(img[0, 1, 0], 0.5 * a[5] + 0.25 * a[3] + 0.25 * a[4], rtol=1e-6)
    np.testing.assert_allclose(img[1, 0, 1], 0.125 * a[0] + 0.625 * a[1] + 0.25 * a[2], rtol=1e-6)
    live = np.zeros((2, 2, 2), bool)
    live[0, 1, 0] = live[1, 0, 1] = True
    assert (img[~live] == 0).all()


def test_triangle_interpolate_reprojects_pixel_centre():
    """Interpolating the clip-space vertices over a perspective raster and
    dividing by the interpolated w gives the NDC of each pixel centre."""
    rng = np.random.default_rng(14)
    H, W = 24, 30
    P = np.zeros((1, 6, 4), np.float32)
    w = rng.uniform(0.6, 2.0, 6)
    ndc = np.array([[-0.8, -0.7], [0.9, -0.5], [0.1, 0.9], [-0.9, 0.8], [0.7, 0.85], [0.0, -0.9]])
    P[0, :, 0] = ndc[:, 0] * w
    P[0, :, 1] = ndc[:, 1] * w
    P[0, :, 2] = np.array([0.1, 0.1, 0.1, 0.4, 0.4, 0.4]) * w
    P[0, :, 3] = w
    tri = np.array([[0, 1, 2], [3, 5, 4]], np.int32)
    rast = oracle.rasterize((H, W), P, tri)
    img = oracle.interpolate_triangle_attributes(rast, P[0], tri).astype(np.float64)
    hit = np.argwhere(rast[0, :, :, 3] > 0)
    assert hit.shape[0] > 100
    for y, x in hit:
        q = img[0, y, x]
        assert abs(q[0] / q[3] - (-1 + 2 * (x + 0.5) / W)) < 2e-5
        assert abs(q[1] / q[3] - (1 - 2 * (y + 0.5) / H)) < 2e-5
        t = int(rast[0, y, x, 3]) - 1
        assert abs(q[2] / q[3] - (0.1 if t == 0 else 0.4)) < 2e-5


def test_triangle_interpolate_backward_hand_built_and_adjoint():
    tri = np.array([[0, 1, 2], [5, 3, 4]], np.int32)
    rast = np.zeros((1, 1, 2, 4), np.float32)
    rast[0, 0, 0] = [0.5, 0.25, 0.0, 2.0]
    rast[0, 0, 1] = [0.125, 0.375, 0.0, 2.0]
    g = np.zeros((1, 1, 2, 1), np.float32)
    g[0, 0, 0] = 2.0
    g[0, 0, 1] = -8.0
    d = oracle.backward_interpolate_triangle_attributes(rast, g, tri, 6)[:, 0].astype(np.float64)
    want = np.zeros(6)
    want[5] = 0.5 * 2.0 + 0.125 * -8.0
    want[3] = 0.25 * 2.0 + 0.375 * -8.0
    want[4] = 0.25 * 2.0 + 0.5 * -8.0
    np.testing.assert_allclose(d, want, rtol=1e-6, atol=0)
    rng = np.random.default_rng(15)
    H, W = 28, 28
    V = 30
    P = clip(rng.uniform(0.0, 28.0, size=(V, 2)), H, W, z=rng.uniform(-0.5, 0.5, size=V))
    tri = rng.integers(0, V, size=(20, 3)).astype(np.int32)
    tri = tri[(tri[:, 0] != tri[:, 1]) & (tri[:, 1] != tri[:, 2]) & (tri[:, 0] != tri[:, 2])]
    rast = oracle.rasterize((H, W), P, tri)
    assert (rast[..., 3] > 0).sum() > 50
    a = rng.normal(size=(V, 2)).astype(np.float32)
    g = rng.normal(size=(1, H, W, 2)).astype(np.float32)
    lhs = (oracle.interpolate_triangle_attributes(rast, a, tri).astype(np.float64) * g).sum()
    rhs = (a.astype(np.float64) *
           oracle.backward_interpolate_triangle_attributes(rast, g, tri, V)).sum()
    scale = np.abs(g).sum() * np.abs(a).max()
    assert abs(lhs - rhs) <= 1e-6 * scale


# --------------------------------------------------------------- arclength


def test_arclength_closed_form():
    """curve_utils.py:167-180: 0 then the running sum of segment lengths;
    integer Pythagorean segments (one of zero length) are exact in fp32."""
    s = np.array([[[0, 0, 0], [3, 4, 0], [3, 4, 0], [3, 4, 12], [0, 0, 0]],
                  [[1, 1, 1], [1, 1, 2], [1, 3, 2], [1, 3, 2], [1, 3, 2]]], np.float32)
    got = oracle.compute_arclength(s)
    np.testing.assert_array_equal(got, np.array([[0, 5, 5, 17, 30], [0, 1, 3, 3, 3]], np.float32))
    rng = np.random.default_rng(16)
    x = rng.normal(size=(7, 33, 3)).astype(np.float32)
    want = np.concatenate(
        [np.zeros((7, 1)),
         np.cumsum(np.linalg.norm(np.diff(x.astype(np.float64), axis=1), axis=2), axis=1)], 1)
    np.testing.assert_allclose(got := oracle.compute_arclength(x), want, rtol=2e-6, atol=1e-6)
    assert got.dtype == np.float32 and (np.diff(got, axis=1) >= 0).all()


def test_absum_bounds_the_value():
    """Every with_abs output is at least the magnitude of its value and zero
    where nothing contributes: smoothing (weights are positive), skinning
    and the antialias gradient."""
    x = np.random.default_rng(17).normal(size=(5, 12, 3)).astype(np.float32)
    s = oracle.compute_arclength(x)
    out, ab = oracle.smooth_strands(x, s, 0.3, with_abs=True)
    assert (ab >= np.abs(out) * (1 - 1e-6)).all() and (ab > 0).all()
    R = np.broadcast_to(np.eye(3, dtype=np.float32), (5, 12, 3, 3)).copy()
    out, ab = oracle.skinning(x, s, x, x, R, 0.3, with_abs=True)
    assert (ab >= np.abs(out) * (1 - 1e-6)).all()
    H = W = 32
    pos = clip([(8.3, 8.3), (24.6, 9.1), (14.2, 25.7)], H, W)
    tri = np.array([[0, 1, 2]], np.int32)
    edges, e2t = np.array([[0, 1], [1, 2], [0, 2]], np.int32), np.array([[0, -1]] * 3, np.int32)
    e2c = np.zeros(3, np.int32)
    cards = oracle.triangle_cluster_mask((H, W), pos, tri, np.zeros(1, np.int32), 1)
    target = np.zeros_like(cards)
    g, ab = oracle.antialiased_cluster_bitset_grad(cards, target, pos, tri, edges, e2t, e2c, 3.0,
                                                   1.0, with_abs=True)
    assert np.abs(g).max() > 0 and (ab >= np.abs(g) * (1 - 1e-6)).all()
    assert (ab[..., 2] == 0).all()  # z carries no gradient


# ------------------------------------------------------------ card fitting loop


def test_cardfit_strip_mesh_counts_and_adjacency():
    """The strip of Q quads: 2Q triangles covering each quad once, 4Q + 1
    unique edges of which the 2Q row edges and the two end w-edges are
    boundary (one triangle) and the Q diagonals and Q - 1 inner w-edges are
    shared by the two triangles on either side; quads are CCW cells."""
    from oracle import cardfit

    for Q in (1, 2, 8, 32):
        N = Q + 1
        tri, quads, edges, e2t, h, w = cardfit.strip_mesh(Q)
        assert tri.shape == (2 * Q, 3) and quads.shape == (Q, 4)
        assert edges.shape == (4 * Q + 1, 2) and (edges[:, 0] < edges[:, 1]).all()
        boundary = {tuple(e) for e, t in zip(edges.tolist(), e2t.tolist()) if t[1] == -1}
        want = {(r * N + j, r * N + j + 1) for r in range(2) for j in range(Q)} | {(0, N), (Q, 2 * Q + 1)}
        assert boundary == want
        for e, (t0, t1) in zip(edges.tolist(), e2t.tolist()):
            assert set(e) <= set(tri[t0].tolist())
            if t1 >= 0:
                assert t1 != t0 and set(e) <= set(tri[t1].tolist())
        for q in range(Q):  # the two triangles of a cell cover its 4 corners
            assert set(tri[2 * q].tolist()) | set(tri[2 * q + 1].tolist()) == set(quads[q].tolist())
        assert {tuple(x) for x in h.tolist()} | {tuple(x) for x in w.tolist()} == \
            {tuple(e) for e in edges.tolist()} - {(q, N + q + 1) for q in range(Q)}


def test_cardfit_regularizer_gradient_is_the_energy_gradient():
    """Central differences of w_pln * sum |q0 + q2 - q1 - q3|^2 +
    w_lap * |L pos|^2, computed directly from the definitions."""
    from oracle import cardfit

    rng = np.random.default_rng(23)
    Q = 5
    _, quads, _, _, h, w = cardfit.strip_mesh(Q)
    V = 2 * (Q + 1)
    L = oracle.laplacian_edges(V, np.concatenate([h, w]))
    pos = rng.normal(size=(V, 3))

    def energy(p):
        pl = sum(np.sum((p[a] + p[c] - p[b] - p[d]) ** 2) for a, b, c, d in quads)
        return 10.0 * pl + 0.01 * np.sum((L @ p) ** 2)

    g = cardfit.regularizer_grad(pos, quads, L, 10.0, 0.01)
    fd = np.zeros_like(pos)
    for i in range(V):
        for k in range(3):
            e = np.zeros_like(pos)
            e[i, k] = 1e-6
            fd[i, k] = (energy(pos + e) - energy(pos - e)) / 2e-6
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_cardfit_clip_chain_rule():
    """d/dpos of sum(gc * clip) with clip = mvp [pos, 1] for a
    non-symmetric mvp, z component of gc dropped."""
    from oracle import cardfit

    rng = np.random.default_rng(29)
    mvp = rng.normal(size=(4, 4))
    gc = rng.normal(size=(6, 4))
    gz = gc.copy()
    gz[:, 2] = 0.0
    pos = rng.normal(size=(6, 3))
    f = lambda p: np.sum(gz * (np.concatenate([p, np.ones((6, 1))], 1) @ mvp.T))  # noqa: E731
    fd = np.zeros_like(pos)
    for i in range(6):
        for k in range(3):
            e = np.zeros_like(pos)
            e[i, k] = 1e-6
            fd[i, k] = (f(pos + e) - f(pos - e)) / 2e-6
    np.testing.assert_allclose(cardfit.clip_grad_to_positions(gc, mvp), fd, rtol=1e-6, atol=1e-7)


def test_cardfit_loop_lowers_the_xor():
    """Twenty iterations of the CPU loop on a small job lower its XOR pixel
    count and keep the vertices finite."""
    from oracle import cardfit
    from strns2cards_amd import fitting

    j = [x for x in fitting.make_jobs(40, seed=3) if x.res == 64][0]
    n, nv = j.strands.shape[:2]
    job = cardfit.job_dict(j.res, j.quads, j.strand_clip.reshape(n, nv, 4), j.mvp, j.verts)
    hyp = cardfit.hyper_dict(fitting.FitHyper())
    _, xor0, tp = cardfit.fit_job(job, hyp, 0)
    pos, xor, tp2 = cardfit.fit_job(job, hyp, 20)
    assert tp == tp2 > 0 and np.isfinite(pos).all()
    assert xor < xor0, (xor0, xor)
