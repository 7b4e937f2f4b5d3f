"""GPU parity: the C ABI (libstrns2cards_hip.so) against the CPU oracle.

Bar: bit-exact for every integer / index / bitset output and for float
outputs computed with the same operation sequence (fragment attributes,
barycentrics, interpolated images).  Float outputs that are sums whose order
differs (atomic scatters) or whose exponentials may differ in the last ulp
(smoothing, skinning), element by element: |got - want| <= 1e-5 |want| +
k u sum|term|, u the fp32 unit roundoff and sum|term| from the oracle
(conftest.elementwise_close).
"""
import numpy as np
import pytest
import torch

import oracle
from capi import call, call_count
from conftest import elementwise_close, report
from strns2cards_amd import synth

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


# ------------------------------------------------------------- inputs


def random_lines(rng, B, V, E, H, W, persp=False):
    """Clip-space vertices spread over and beyond the screen, with
    degenerate, axis-aligned and off-screen edges mixed in."""
    xy = rng.uniform(-1.3, 1.3, size=(B, V, 2))
    z = rng.uniform(-1.2, 1.2, size=(B, V, 1))
    w = rng.uniform(0.5, 2.0, size=(B, V, 1)) if persp else np.ones((B, V, 1))
    pos = np.concatenate([xy * w, z * w, w], -1).astype(np.float32)
    edges = rng.integers(0, V, size=(E, 2)).astype(np.int32)
    edges[: E // 10, 1] = edges[: E // 10, 0]  # zero-length
    k = E // 10
    pos[:, edges[k:2 * k, 1], 1] = pos[:, edges[k:2 * k, 0], 1]  # horizontal in NDC (w = 1)
    return pos, edges


def dda_gpu(dev, res, pos, edges, edge_ids=None):
    H, W = res
    B, V, _ = pos.shape
    E = edges.shape[0]
    ids = np.arange(B * E, dtype=np.int32) if edge_ids is None else edge_ids
    tp, te, ti = T(pos, dev), T(edges, dev), T(ids, dev)
    P = ids.shape[0]
    prefix = torch.empty(P, dtype=torch.int32, device=dev)
    slopes = torch.full((P, 2), float("nan"), device=dev)
    spans = torch.full((P, 4), float("nan"), device=dev)
    n = call_count("s2c_dda_compute_span", P, ti, H, W, V, tp, E, te, prefix, slopes, spans)
    fp = torch.full((n, 3), -1, dtype=torch.int32, device=dev)
    fa = torch.zeros((n, 4), device=dev)
    call("s2c_dda_compute_fragments", P, n, prefix, ti, slopes, spans, H, W, V, tp, E, te, fp, fa)
    fpo, fao = torch.empty_like(fp), torch.empty_like(fa)
    m = call_count("s2c_filter_valid_fragments", n, fp, fa, fpo, fao)
    # the survivor-space form (count, then emit) gives the same list, with
    # the ids given and, when they are every edge in order, with null ids
    for ids_arg in ([ti] + ([None] if edge_ids is None else [])):
        vp = torch.full((P,), -9, dtype=torch.int32, device=dev)
        mc = call_count("s2c_dda_count_fragments", P, ids_arg, H, W, V, tp, E, te, vp)
        assert mc == m
        fpc = torch.full((m, 3), 7, dtype=torch.int32, device=dev)
        fac = torch.full((m, 4), 7.0, device=dev)
        call("s2c_dda_emit_fragments", P, m, vp, ids_arg, H, W, V, tp, E, te, fpc, fac)
        assert torch.equal(fpc, fpo[:m]) and torch.equal(fac, fao[:m])
    torch.cuda.synchronize()
    return (prefix.cpu().numpy(), slopes.cpu().numpy(), spans.cpu().numpy(), n,
            fpo[:m].cpu().numpy(), fao[:m].cpu().numpy())


def check_dda(dev, res, pos, edges, edge_ids=None):
    prefix, slopes, spans, n, fp, fa = dda_gpu(dev, res, pos, edges, edge_ids)
    o_prefix, o_slopes, o_spans, o_n = oracle.dda_compute_span(res, pos, edges, edge_ids)
    assert n == o_n
    np.testing.assert_array_equal(prefix, o_prefix)
    np.testing.assert_array_equal(slopes, o_slopes)  # NaN == NaN where unwritten
    np.testing.assert_array_equal(spans, o_spans)
    o_fp, o_fa = oracle.dda_compute_fragments(res, pos, edges, edge_ids)
    np.testing.assert_array_equal(fp, o_fp)
    np.testing.assert_array_equal(fa, o_fa)
    return fp, fa


# ------------------------------------------------------------------ DDA


@pytest.mark.parametrize("persp", [False, True])
@pytest.mark.parametrize("res", [(37, 53), (128, 96)])
def test_dda_random_lines(dev, res, persp):
    report(f"dda random lines res={res} persp={persp}")
    rng = np.random.default_rng(11 + res[0] + persp)
    pos, edges = random_lines(rng, 3, 400, 3000, *res, persp=persp)
    check_dda(dev, res, pos, edges)
    # subset of primitives in arbitrary order, as the antialias path uses
    ids = rng.permutation(3 * 3000)[:1000].astype(np.int32)
    check_dda(dev, res, pos, edges, ids)


def test_dda_small_workload(dev):
    report("dda small workload (2k strands, 256^2)")
    w = synth.small_workload()
    fp, fa = check_dda(dev, (w.H, w.W), w.pos, w.edges)
    assert fp.shape[0] == 133424


def test_dda_empty(dev):
    report("dda empty inputs")
    pos = np.zeros((1, 4, 4), np.float32)
    pos[..., 3] = 1
    edges = np.zeros((0, 2), np.int32)
    prefix, slopes, spans, n, fp, fa = dda_gpu(dev, (8, 8), pos, edges)
    assert n == 0 and fp.shape == (0, 3)


# --------------------------------------------------------- DDA -> mask


def fused_dda_mask(dev, res, pos, edges, e2c, K, tile_log2=0):
    S = (K + 31) // 32
    B, V, _ = pos.shape
    bits = torch.full((B, res[0], res[1], S), 0x5A5A5A5A, dtype=torch.int32, device=dev)
    if tile_log2:
        call("s2c_dda_cluster_mask_tiled", B, res[0], res[1], V, T(pos, dev), edges.shape[0],
             T(edges, dev), T(e2c, dev), S, tile_log2, bits)
    else:
        call("s2c_dda_cluster_mask", B, res[0], res[1], V, T(pos, dev), edges.shape[0],
             T(edges, dev), T(e2c, dev), S, bits)
    return bits.cpu().numpy()


def test_dda_mask_small_workload(dev):
    report("dda->mask small workload")
    w = synth.small_workload()
    want = oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, 64)
    got = fused_dda_mask(dev, (w.H, w.W), w.pos, w.edges, w.edge2cluster, 64)
    np.testing.assert_array_equal(got, want)
    # and the list form: fragments -> cluster_mask_from_fragments
    _, _, _, _, fp, fa = dda_gpu(dev, (w.H, w.W), w.pos, w.edges)
    bits = torch.zeros((1, w.H, w.W, 2), dtype=torch.int32, device=dev)
    call("s2c_cluster_mask_from_fragments", fp.shape[0], T(fp, dev), T(fa, dev),
         T(w.edge2cluster, dev), w.H, w.W, 2, bits)
    np.testing.assert_array_equal(bits.cpu().numpy(), want)


@pytest.mark.parametrize("K", [1, 33, 1024, 17000])  # 17000: tile bitset exceeds LDS -> direct ORs
def test_dda_mask_random(dev, K):
    report(f"dda->mask random K={K}")
    rng = np.random.default_rng(K)
    pos, edges = random_lines(rng, 2, 300, 2500, 61, 70, persp=True)
    e2c = rng.integers(0, K, size=edges.shape[0]).astype(np.int32)
    want = oracle.dda_cluster_mask((61, 70), pos, edges, e2c, K)
    got = fused_dda_mask(dev, (61, 70), pos, edges, e2c, K)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("B,E", [(37, 100), (9, 1), (5, 3), (3, 1023), (3, 1025)])
def test_dda_mask_short_views(dev, B, E):
    """Views with fewer edges than a workgroup walks (1,024): one workgroup
    spans several views, including views of one edge, and the view changes
    fall at and around workgroup boundaries."""
    report(f"dda->mask short views B={B} E={E}")
    rng = np.random.default_rng(B * 7919 + E)
    pos, edges = random_lines(rng, B, 50, E, 40, 52, persp=True)
    e2c = rng.integers(0, 70, size=E).astype(np.int32)
    want = oracle.dda_cluster_mask((40, 52), pos, edges, e2c, 70)
    got = fused_dda_mask(dev, (40, 52), pos, edges, e2c, 70)
    np.testing.assert_array_equal(got, want)


def test_dda_mask_long_scattered_edges(dev):
    """Long random edges over 1024^2: deferred (workgroup-walked) remainders,
    and workgroups touching more distinct tiles than their LDS table holds
    (the global overflow path)."""
    report("dda->mask long scattered edges (overflow + deferred paths)")
    rng = np.random.default_rng(31)
    pos, edges = random_lines(rng, 2, 6000, 5000, 1024, 1024, persp=False)
    e2c = rng.integers(0, 100, size=edges.shape[0]).astype(np.int32)
    want = oracle.dda_cluster_mask((1024, 1024), pos, edges, e2c, 100, threads=8)
    got = fused_dda_mask(dev, (1024, 1024), pos, edges, e2c, 100)
    np.testing.assert_array_equal(got, want)


def test_dda_mask_table_overflow_twice(dev):
    """Workgroups of long lines (every one past its key table: their
    fragments are OR-ed behind the binned pass) alternating with workgroups
    of short ones (binned), twice in a row on the same buffers; both masks
    equal the oracle's."""
    report("dda->mask key-table overflow mixed with binned workgroups, twice")
    rng = np.random.default_rng(37)
    pos_l, edges_l = random_lines(rng, 1, 3000, 2048, 1024, 1024, persp=False)
    pos_s, edges_s = _segments(rng, 2048, 1.5, 3.0, 160, 1024, 1024)
    pos = np.concatenate([pos_l, pos_s], 1)
    edges_s = edges_s + pos_l.shape[1]
    edges = np.concatenate([edges_l[:1024], edges_s[:1024], edges_l[1024:], edges_s[1024:]])
    e2c = rng.integers(0, 100, size=edges.shape[0]).astype(np.int32)
    e2c[1024:2048] %= 4  # the short lines' workgroups: few clusters
    e2c[3072:] %= 4
    keys = _workgroup_keys(pos, edges, e2c, (1024, 1024))
    assert keys[0] > 512 and keys[2] > 512 and keys[1] < 400 and keys[3] < 400
    want = oracle.dda_cluster_mask((1024, 1024), pos, edges, e2c, 100, threads=8)
    for _ in range(2):
        got = fused_dda_mask(dev, (1024, 1024), pos, edges, e2c, 100)
        np.testing.assert_array_equal(got, want)


def _segments(rng, n, lo, hi, region, H, W):
    """n segments of length in [lo, hi] px with random starts and directions
    inside the top-left region x region pixels of an H x W image (one view,
    w = 1)."""
    p0 = rng.uniform(0, region, size=(n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    p1 = np.clip(p0 + np.stack([np.cos(ang), np.sin(ang)], 1) * rng.uniform(lo, hi, n)[:, None],
                 0, region - 0.01)
    pts = np.stack([p0, p1], 1).reshape(-1, 2)
    pos = np.zeros((1, 2 * n, 4), np.float32)
    pos[0, :, 0] = 2 * pts[:, 0] / W - 1
    pos[0, :, 1] = 1 - 2 * pts[:, 1] / H
    pos[0, :, 3] = 1
    return pos, np.arange(2 * n, dtype=np.int32).reshape(n, 2)


def _workgroup_keys(pos, edges, e2c, res, cell=16):
    """Distinct (16 x 16 cell, cluster) keys of each 1,024-edge workgroup
    of the fused pass over view 0 (the oracle's fragments)."""
    fp, fa = oracle.dda_compute_fragments(res, pos[:1], edges)
    eid = fa[:, 3].astype(np.int64) - 1
    k = eid >= 0
    key = (((eid[k] // 1024) * 4096 + fp[k, 1] // cell) * 4096 + fp[k, 2] // cell) * 65536 \
        + e2c[eid[k]]
    return np.bincount(np.unique(key) // (4096 * 4096 * 65536))


@pytest.mark.parametrize("case", ["near_capacity", "table_overflow", "cells_in_large_tiles"])
def test_dda_mask_workgroup_paths(dev, case):
    """Workgroups (1,024 edges each) with 370-390 distinct (cell, cluster)
    keys in their 512-slot table (long probe sequences); with about 1,170
    (the table overflows: fragments OR-ed behind the binned pass); and
    64-pixel tiles of 4 x 4 cells (40-60 pixel edges inside one tile, 16
    clusters), so records carry the cell's place in the tile."""
    report(f"dda->mask workgroup path: {case}")
    rng = np.random.default_rng(5)
    lo, hi, region, K = {"near_capacity": (1.5, 3.0, 160, 4),
                         "table_overflow": (1.5, 3.0, 448, 64),
                         "cells_in_large_tiles": (40.0, 60.0, 63, 16)}[case]
    pos, edges = _segments(rng, 4096, lo, hi, region, 1024, 1024)
    e2c = rng.integers(0, K, size=edges.shape[0]).astype(np.int32)
    keys = _workgroup_keys(pos, edges, e2c, (1024, 1024))
    lo_k, hi_k = {"near_capacity": (350, 420), "table_overflow": (1000, 4096),
                  "cells_in_large_tiles": (200, 300)}[case]
    assert lo_k <= keys.min() and keys.max() <= hi_k, keys
    want = oracle.dda_cluster_mask((1024, 1024), pos, edges, e2c, 64, threads=8)
    got = fused_dda_mask(dev, (1024, 1024), pos, edges, e2c, 64,
                         tile_log2=6 if case == "cells_in_large_tiles" else 0)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("tile_log2", [2, 3])
def test_dda_mask_many_tiles(dev, tile_log2):
    """Screen tiles of 4 x 4 and 8 x 8 pixels at 1040 x 1024: 66,560 and
    16,640 tiles, more than the counting sort's LDS counters hold at once
    (16,384), so its count and placement passes take several rounds; the
    cells are the tiles, with 1- and 2-word masks."""
    report(f"dda->mask many tiles, tile side {1 << tile_log2}")
    rng = np.random.default_rng(61 + tile_log2)
    res = (1040, 1024)
    pos, edges = random_lines(rng, 1, 3000, 6000, *res, persp=True)
    seg_pos, seg_edges = _segments(rng, 4096, 2.0, 40.0, 1000, *res)
    pos = np.concatenate([pos, seg_pos], 1)
    edges = np.concatenate([edges, seg_edges + 3000])
    e2c = rng.integers(0, 32, size=edges.shape[0]).astype(np.int32)
    want = oracle.dda_cluster_mask(res, pos, edges, e2c, 32, threads=8)
    got = fused_dda_mask(dev, res, pos, edges, e2c, 32, tile_log2=tile_log2)
    np.testing.assert_array_equal(got, want)


def test_dda_mask_keys_beyond_24_bits(dev):
    """Ten views of the same 20 short segments (one workgroup of 200 edges)
    at 64 x 64 with 16,384 clusters: 4 x 4 cells, 14 cluster bits, so a key
    is ((b * 16 + cell row) << 4 | cell column) << 14 | cluster and the keys
    of views b and b + 4 differ only above bit 23.  The key table's hash
    reads the low 24 bits, so those keys share their home slot and are
    placed by probing; the mask equals the oracle's."""
    report("dda->mask keys that differ only above bit 23")
    rng = np.random.default_rng(73)
    res, K, B = (64, 64), 16384, 10
    pos1, edges = _segments(rng, 20, 1.5, 3.0, 60, *res)
    pos = np.repeat(pos1, B, axis=0)
    e2c = rng.integers(0, K, size=edges.shape[0]).astype(np.int32)
    fp, fa = oracle.dda_compute_fragments(res, pos, edges)
    eid = fa[:, 3].astype(np.int64) - 1
    ok = eid >= 0
    key = ((((fp[ok, 0].astype(np.int64) * 16 + fp[ok, 1] // 4) << 4) | (fp[ok, 2] // 4)) << 14) \
        | e2c[eid[ok] % edges.shape[0]]
    key = np.unique(key)
    low, n = np.unique(key & 0xFFFFFF, return_counts=True)
    assert key.size <= 400 and key.max() >= 1 << 24, key.size
    assert (n >= 2).sum() * 3 >= low.size, (n >= 2).sum()  # most home slots shared
    want = oracle.dda_cluster_mask(res, pos, edges, e2c, K, threads=8)
    got = fused_dda_mask(dev, res, pos, edges, e2c, K)
    np.testing.assert_array_equal(got, want)


def test_dda_mask_nonfinite_and_huge_vertices(dev):
    """Vertices at +-inf, NaN, 1e30 and w = 0 beside ordinary ones.  An
    infinite x or y makes dda.cu:124-144's t1 NaN, so those fragments fail
    the depth test even where z and w alone would pass it; edges reaching far
    off screen are long, and workgroups mix them with ordinary edges."""
    report("dda->mask non-finite and huge vertices")
    rng = np.random.default_rng(41)
    pos, edges = random_lines(rng, 2, 400, 3000, 96, 80, persp=True)
    bad = rng.choice(400, size=60, replace=False)
    vals = np.array([np.inf, -np.inf, np.nan, 1e30, -1e30, 3e4], np.float32)
    for j, v in enumerate(bad):
        pos[j % 2, v, rng.integers(0, 2)] = vals[j % len(vals)]
    pos[0, bad[:6], 3] = 0.0  # w = 0: the NDC divide overflows
    e2c = rng.integers(0, 50, size=edges.shape[0]).astype(np.int32)
    want = oracle.dda_cluster_mask((96, 80), pos, edges, e2c, 50)
    got = fused_dda_mask(dev, (96, 80), pos, edges, e2c, 50)
    np.testing.assert_array_equal(got, want)


def test_dda_mask_depth_boundaries(dev):
    """Edges whose endpoint depths sit at / around the fast-path threshold
    (|z_ndc| = 1 - 1e-5) and at the clip limit, with w = 1 and with w
    ratios beyond 16 (which must take the exact per-fragment test)."""
    report("dda->mask depth boundaries")
    rng = np.random.default_rng(21)
    n = 4000
    pos = np.zeros((2, 2 * n, 4), np.float32)
    zs = np.array([1 - 1e-5, 1 - 2e-5, 1 - 5e-6, 1.0, 1 + 1e-7, 0.0, 0.999999], np.float32)
    xy = rng.uniform(-1.1, 1.1, size=(2, 2 * n, 2)).astype(np.float32)
    z = rng.choice(np.concatenate([zs, -zs]), size=(2, 2 * n)).astype(np.float32)
    w = np.where(rng.uniform(size=(2, 2 * n)) < 0.5, 1.0,
                 rng.uniform(0.05, 3.0, size=(2, 2 * n))).astype(np.float32)
    pos[..., 0], pos[..., 1], pos[..., 2], pos[..., 3] = xy[..., 0] * w, xy[..., 1] * w, z * w, w
    edges = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.int32)
    e2c = rng.integers(0, 40, size=n).astype(np.int32)
    want = oracle.dda_cluster_mask((90, 70), pos, edges, e2c, 40)
    got = fused_dda_mask(dev, (90, 70), pos, edges, e2c, 40)
    np.testing.assert_array_equal(got, want)


def list_mask(dev, B, res, fp, fa, p2c, K):
    """s2c_fragment_cluster_mask over a fragment list (every word written)."""
    S = (K + 31) // 32
    bits = torch.full((B, res[0], res[1], S), 0x5A5A5A5A, dtype=torch.int32, device=dev)
    call("s2c_fragment_cluster_mask", B, res[0], res[1], fp.shape[0], T(fp, dev), T(fa, dev),
         T(p2c, dev), S, bits)
    return bits.cpu().numpy()


@pytest.mark.parametrize("K", [1, 33, 1024, 17000])  # 17000: tile bitset exceeds LDS -> direct ORs
def test_fragment_mask_lists(dev, K):
    """The binned mask of a fragment list (rasterize.py:170-214) against the
    oracle's cluster_mask_from_fragments, for DDA and triangle fragments."""
    report(f"fragment-list mask K={K}")
    rng = np.random.default_rng(K + 5)
    res = (61, 70)
    pos, edges = random_lines(rng, 2, 300, 2500, *res, persp=True)
    fp, fa = oracle.dda_compute_fragments(res, pos, edges)
    e2c = rng.integers(0, K, size=edges.shape[0]).astype(np.int32)
    want = oracle.cluster_mask_from_fragments(res, fp, fa, e2c, K, B=2)
    np.testing.assert_array_equal(list_mask(dev, 2, res, fp, fa, e2c, K), want)
    tpos = random_tris(rng, 2, 200, 300, persp=True)[0]
    tri = rng.integers(0, 200, size=(300, 3)).astype(np.int32)
    tfp, tfa = oracle.compute_fragments(res, tpos, tri)
    t2c = rng.integers(0, K, size=300).astype(np.int32)
    want = oracle.cluster_mask_from_fragments(res, tfp, tfa, t2c, K, B=2)
    np.testing.assert_array_equal(list_mask(dev, 2, res, tfp, tfa, t2c, K), want)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 4097, 300001])
def test_max_frag_view(dev, n):
    """s2c_max_frag_view: the largest view index of a fragment list, for
    lengths around its 4-row steps."""
    rng = np.random.default_rng(n)
    fp = np.stack([rng.integers(-5, 40, n), rng.integers(0, 9, n), rng.integers(0, 9, n)],
                  1).astype(np.int32)
    fp[rng.integers(0, n), 0] = 57
    m = call_count("s2c_max_frag_view", n, T(fp, dev))
    assert m == fp[:, 0].max() == 57


def test_fragment_mask_invalid_entries(dev):
    """Prim ids <= 0, pixels outside the image and clusters beyond the
    bitset set no bit; the other fragments set theirs."""
    report("fragment-list mask with invalid entries")
    rng = np.random.default_rng(77)
    res, B, K, n = (40, 52), 3, 70, 20000
    fp = np.stack([rng.integers(0, B, n), rng.integers(0, res[0], n),
                   rng.integers(0, res[1], n)], 1).astype(np.int32)
    fa = rng.uniform(size=(n, 4)).astype(np.float32)
    fa[:, 3] = rng.integers(1, 500, n)
    p2c = rng.integers(0, K, size=500).astype(np.int32)
    bad = rng.choice(n, size=3000, replace=False)
    fa[bad[:600], 3] = 0.0
    fa[bad[600:900], 3] = -3.0
    fp[bad[900:1300], 0] = B
    fp[bad[1300:1700], 1] = -1
    fp[bad[1700:2100], 2] = res[1]
    p2c_bad = p2c.copy()
    p2c_bad[:20] = 96  # >= 32 * S (S = 3)
    ok = np.ones(n, bool)
    ok[bad[:2100]] = False
    prim = fa[:, 3].astype(np.int64) - 1
    ok &= ~np.isin(prim, np.arange(20))
    want = oracle.cluster_mask_from_fragments(res, fp[ok], fa[ok], p2c_bad, K, B=B)
    np.testing.assert_array_equal(list_mask(dev, B, res, fp, fa, p2c_bad, K), want)


def test_fragment_mask_table_overflow_twice(dev):
    """Fragments at random pixels of 4 views at 1024^2: every workgroup's
    4,096 fragments have more distinct (cell, cluster) keys than its table
    holds, so all are OR-ed behind the binned pass; twice in a row."""
    report("fragment-list mask key-table overflow, twice")
    rng = np.random.default_rng(91)
    res, B, K, n = (1024, 1024), 4, 256, 300000
    fp = np.stack([rng.integers(0, B, n), rng.integers(0, res[0], n),
                   rng.integers(0, res[1], n)], 1).astype(np.int32)
    fa = np.zeros((n, 4), np.float32)
    fa[:, 3] = rng.integers(1, 1000, n)
    p2c = rng.integers(0, K, size=1000).astype(np.int32)
    want = oracle.cluster_mask_from_fragments(res, fp, fa, p2c, K, B=B)
    for _ in range(2):
        np.testing.assert_array_equal(list_mask(dev, B, res, fp, fa, p2c, K), want)


def test_fragment_mask_keys_beyond_24_bits(dev):
    """The fragment-list form of test_dda_mask_keys_beyond_24_bits: the ten
    views' fragments interleaved (fragment i of each view in turn), all in
    one 4,096-fragment workgroup, so keys that differ only above bit 23
    share home slots of its table; the mask equals the oracle's."""
    report("fragment-list mask keys that differ only above bit 23")
    rng = np.random.default_rng(73)
    res, K, B = (64, 64), 16384, 10
    pos1, edges = _segments(rng, 20, 1.5, 3.0, 60, *res)
    pos = np.repeat(pos1, B, axis=0)
    e2c = rng.integers(0, K, size=edges.shape[0]).astype(np.int32)
    fp, fa = oracle.dda_compute_fragments(res, pos, edges)
    n = fp.shape[0]
    assert n % B == 0 and n <= 4096, n  # the same fragments in every view
    fp = np.ascontiguousarray(fp.reshape(B, n // B, 3).transpose(1, 0, 2).reshape(n, 3))
    fa = np.ascontiguousarray(fa.reshape(B, n // B, 4).transpose(1, 0, 2).reshape(n, 4))
    assert (fp[:B, 0] == np.arange(B)).all()
    want = oracle.cluster_mask_from_fragments(res, fp, fa, e2c, K, B=B)
    np.testing.assert_array_equal(list_mask(dev, B, res, fp, fa, e2c, K), want)


@pytest.mark.slow
def test_dda_mask_full_workload(dev):
    report("dda->mask full workload (50k strands x 100, 4 views 1024^2, K=1024)")
    w = synth.full_workload()
    want = oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, 1024, threads=16)
    got = fused_dda_mask(dev, (w.H, w.W), w.pos, w.edges, w.edge2cluster, 1024)
    np.testing.assert_array_equal(got, want)
    # the drop-in pair (dda_compute_fragments, cluster_mask_from_fragments)
    import strns2cards_amd as s2c
    pos, edges = T(w.pos, dev), T(w.edges, dev)
    fp, fa = s2c.dda_compute_fragments((w.H, w.W), pos, edges)
    got = s2c.cluster_mask_from_fragments((w.H, w.W), fp, fa, T(w.edge2cluster, dev), 1024)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


# ------------------------------------------------------------ triangles


def tri_gpu(dev, res, pos, tri):
    H, W = res
    B, V, _ = pos.shape
    Tn = tri.shape[0]
    tp, tt = T(pos, dev), T(tri, dev)
    rects = torch.empty((B * Tn, 4), dtype=torch.int32, device=dev)
    prefix = torch.empty(B * Tn, dtype=torch.int32, device=dev)
    n = call_count("s2c_compute_triangle_rects", H, W, B, V, tp, Tn, tt, rects, prefix)
    fp = torch.full((n, 3), -1, dtype=torch.int32, device=dev)
    fa = torch.zeros((n, 4), device=dev)
    call("s2c_compute_fragments", H, W, V, tp, Tn, tt, B * Tn, n, prefix, rects, fp, fa)
    fpo, fao = torch.empty_like(fp), torch.empty_like(fa)
    m = call_count("s2c_filter_valid_fragments", n, fp, fa, fpo, fao)
    return rects.cpu().numpy(), prefix.cpu().numpy(), n, fpo[:m], fao[:m]


def random_tris(rng, B, V, Tn, persp):
    xy = rng.uniform(-1.2, 1.2, size=(B, V, 2))
    z = rng.uniform(-1.1, 1.1, size=(B, V, 1))
    w = rng.uniform(0.6, 1.8, size=(B, V, 1)) if persp else np.ones((B, V, 1))
    pos = np.concatenate([xy * w, z * w, w], -1).astype(np.float32)
    tri = rng.integers(0, V, size=(Tn, 3)).astype(np.int32)
    return pos, tri


@pytest.mark.parametrize("persp", [False, True])
def test_triangles_random(dev, persp):
    report(f"triangles random persp={persp}")
    rng = np.random.default_rng(5 + persp)
    res = (45, 64)
    pos, tri = random_tris(rng, 2, 60, 150, persp)
    rects, prefix, n, fp, fa = tri_gpu(dev, res, pos, tri)
    o_rects, o_prefix, o_n = oracle.compute_triangle_rects(res, pos, tri)
    assert n == o_n
    np.testing.assert_array_equal(rects, o_rects)
    np.testing.assert_array_equal(prefix, o_prefix)
    o_fp, o_fa = oracle.compute_fragments(res, pos, tri)
    np.testing.assert_array_equal(fp.cpu().numpy(), o_fp)
    np.testing.assert_array_equal(fa.cpu().numpy(), o_fa)
    # depth test
    rast = torch.zeros((2, *res, 4), device=dev)
    call("s2c_depth_test", 2, res[0], res[1], fp.shape[0], fp, fa, rast)
    o_rast = oracle.depth_test(2, res, o_fp, o_fa)
    np.testing.assert_array_equal(rast.cpu().numpy(), o_rast)
    # interpolation forward (same op sequence -> exact) and backward (atomics)
    C = 5
    attr = rng.normal(size=(60, C)).astype(np.float32)
    img = torch.zeros((2, *res, C), device=dev)
    call("s2c_interpolate_triangle_attributes", 2, res[0], res[1], C, rast, T(attr, dev), T(tri, dev),
         img)
    np.testing.assert_array_equal(img.cpu().numpy(),
                                  oracle.interpolate_triangle_attributes(o_rast, attr, tri))
    d_img = rng.normal(size=(2, *res, C)).astype(np.float32)
    d_attr = torch.zeros((60, C), device=dev)
    call("s2c_backward_interpolate_triangle_attributes", 2, res[0], res[1], C, rast, T(d_img, dev),
         T(tri, dev), d_attr)
    want, ab = oracle.backward_interpolate_triangle_attributes(o_rast, d_img, tri, 60, True)
    elementwise_close(d_attr.cpu().numpy(), want, ab, what="triangle interpolation backward")


def test_triangles_card_workload(dev):
    report("triangles: small workload cards -> fragments, mask, fused mask")
    w = synth.small_workload()
    c = w.cards
    res = (w.H, w.W)
    _, _, _, fp, fa = tri_gpu(dev, res, w.card_pos, c.tri)
    o_fp, o_fa = oracle.compute_fragments(res, w.card_pos, c.tri)
    np.testing.assert_array_equal(fp.cpu().numpy(), o_fp)
    np.testing.assert_array_equal(fa.cpu().numpy(), o_fa)
    want = oracle.triangle_cluster_mask(res, w.card_pos, c.tri, c.tri2cluster, 64)
    bits = torch.full((1, *res, 2), -1, dtype=torch.int32, device=dev)
    call("s2c_triangle_cluster_mask", 1, res[0], res[1], w.card_pos.shape[1], T(w.card_pos, dev),
         c.tri.shape[0], T(c.tri, dev), T(c.tri2cluster, dev), 2, bits)
    np.testing.assert_array_equal(bits.cpu().numpy(), want)


def fused_tri_mask(dev, res, pos, tri, t2c, K, fill=-1):
    """s2c_triangle_cluster_mask into a bitset prefilled with `fill` (the
    tiled path writes every word; the direct path clears first)."""
    B, V, _ = pos.shape
    S = (K + 31) // 32
    bits = torch.full((B, *res, S), fill, dtype=torch.int32, device=dev)
    call("s2c_triangle_cluster_mask", B, res[0], res[1], V, T(pos, dev), tri.shape[0],
         T(tri, dev), T(t2c, dev), S, bits)
    return bits.cpu().numpy()


# K=1..1024: 16x16 tiles; 2048: 8x8 tiles; 16384: 4x4 tiles; 17000: direct ORs
@pytest.mark.parametrize("K", [1, 33, 1024, 2048, 16384, 17000])
def test_triangle_mask_tiles(dev, K):
    """Tiled triangle -> bitset against the oracle: three perspective views
    of random triangles over a ragged 70 x 45 image (partial edge tiles),
    with screen-filling, off-screen, degenerate and z-clipped triangles."""
    report(f"triangle mask tiles K={K}")
    rng = np.random.default_rng(K + 3)
    pos, tri = random_tris(rng, 3, 90, 400, persp=True)
    pos[:, :3] = np.array([[-3, -3, 0, 1], [3, -3, 0, 1], [0, 3, 0, 1]], np.float32)
    tri[0] = (0, 1, 2)  # covers the whole screen: every tile lists it
    tri[1] = (5, 5, 6)  # degenerate
    t2c = rng.integers(0, K, size=tri.shape[0]).astype(np.int32)
    want = oracle.triangle_cluster_mask((45, 70), pos, tri, t2c, K)
    got = fused_tri_mask(dev, (45, 70), pos, tri, t2c, K)
    np.testing.assert_array_equal(got, want)


@pytest.mark.slow
def test_triangle_mask_full_cards(dev):
    report("triangle mask full workload cards (1024 x 20 quads, 4 views 1024^2, K=1024)")
    w = synth.full_workload()
    c = w.cards
    want = oracle.triangle_cluster_mask((w.H, w.W), w.card_pos, c.tri, c.tri2cluster, 1024)
    got = fused_tri_mask(dev, (w.H, w.W), w.card_pos, c.tri, c.tri2cluster, 1024)
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- bitset


@pytest.mark.parametrize("S", [1, 3, 4, 32])
def test_bitset_reductions(dev, S):
    report(f"bitset rgb/popcount S={S}")
    rng = np.random.default_rng(S)
    bits = rng.integers(-2**31, 2**31, size=(2, 19, 23, S), dtype=np.int64).astype(np.int32)
    bits[bits % 3 == 0] = 0
    rgb = rng.uniform(size=(32 * S, 3)).astype(np.float32)
    out = torch.zeros((2, 19, 23, 3), device=dev)
    call("s2c_accumulate_bitset_rgb", 2, 19, 23, S, T(bits, dev), T(rgb, dev), out)
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.accumulate_bitset_rgb(bits, rgb))
    cnt = torch.zeros((2, 19, 23), device=dev)
    call("s2c_popcount_bitset", 2, 19, 23, S, T(bits, dev), cnt)
    np.testing.assert_array_equal(cnt.cpu().numpy(), oracle.popcount_bitset(bits))


# ------------------------------------------------------------- antialias


def test_antialias_small_workload(dev):
    report("antialias: mark, compact, silhouette DDA, backward (small workload)")
    w = synth.small_workload()
    c = w.cards
    res = (w.H, w.W)
    B, V, _ = w.card_pos.shape
    E = c.edges.shape[0]
    target = oracle.dda_cluster_mask(res, w.pos, w.edges, w.edge2cluster, 64)
    cards = oracle.triangle_cluster_mask(res, w.card_pos, c.tri, c.tri2cluster, 64)
    tp = T(w.card_pos, dev)
    prim = torch.full((B * E,), -1, dtype=torch.int32, device=dev)
    normals = torch.zeros((B * E, 2), device=dev)
    call("s2c_mark_discontinuity_edges", res[0], res[1], B, V, tp, T(c.tri, dev), E, T(c.edges, dev),
         T(c.edge2tri, dev), prim, normals)
    out = torch.empty_like(prim)
    m = call_count("s2c_compact_valid_ints", B * E, prim, out)
    o_prim, o_normals = oracle.mark_discontinuity_edges(res, w.card_pos, c.tri, c.edges, c.edge2tri)
    np.testing.assert_array_equal(out[:m].cpu().numpy(), o_prim)
    np.testing.assert_array_equal(normals.cpu().numpy(), o_normals)
    _, _, _, _, fp, fa = dda_gpu(dev, res, w.card_pos, c.edges, o_prim)
    check_antialias_backward(dev, w, cards, target, fp, fa, normals)


def check_antialias_backward(dev, w, cards, target, fp, fa, normals, edges=None, e2t=None, e2c=None):
    c = w.cards
    edges = c.edges if edges is None else edges
    e2t = c.edge2tri if e2t is None else e2t
    e2c = c.edge2cluster if e2c is None else e2c
    res = (w.H, w.W)
    B, V, _ = w.card_pos.shape
    E = edges.shape[0]
    tp = T(w.card_pos, dev)
    for rho in (1.0, 0.5):
        d_pos = torch.zeros((B, V, 4), device=dev)
        call("s2c_backward_antialiased_cluster_bitset", fp.shape[0], T(fp, dev), T(fa, dev), res[0],
             res[1], 2, T(cards, dev), T(target, dev), V, tp, E, T(edges, dev), normals,
             T(e2c, dev), d_pos, 7.0, rho)
        want, ab = oracle.antialiased_cluster_bitset_grad(cards, target, w.card_pos, c.tri, edges,
                                                          e2t, e2c, 7.0, rho, True)
        assert np.abs(want).max() > 0
        elementwise_close(d_pos.cpu().numpy(), want, ab, what=f"antialias backward rho={rho}")


def test_antialias_shuffled_edges(dev):
    """Edges in random order (as an unordered edge map produces): waves mix
    clusters, so the unstaged window path runs."""
    report("antialias backward with shuffled card edges")
    w = synth.small_workload()
    c = w.cards
    res = (w.H, w.W)
    perm = np.random.default_rng(4).permutation(c.edges.shape[0])
    edges, e2t, e2c = c.edges[perm].copy(), c.edge2tri[perm].copy(), c.edge2cluster[perm].copy()
    target = oracle.dda_cluster_mask(res, w.pos, w.edges, w.edge2cluster, 64)
    cards = oracle.triangle_cluster_mask(res, w.card_pos, c.tri, c.tri2cluster, 64)
    prim, normals = oracle.mark_discontinuity_edges(res, w.card_pos, c.tri, edges, e2t)
    fp, fa = oracle.dda_compute_fragments(res, w.card_pos, edges, prim)
    check_antialias_backward(dev, w, cards, target, fp, fa, T(normals, dev), edges, e2t, e2c)


@pytest.mark.parametrize("mult", [2, 8, 16, 17])  # S = 4, 16, 32, 34 words per pixel
def test_antialias_word_planes(dev, mult):
    """Antialias backward through the entry point that is given the view
    count, which reads the residual bits from word planes (a transpose of
    both bitsets) when S >= 16 and the pixel records below, and through the
    reference-shaped one, which always reads the records.  Two views of the
    small workload (the second shifted in clip space); cluster ids spread as
    mult*k + 1 so S = 4 (records), 16 (the threshold), 32 and 34 (a partial
    second chunk of 32 words)."""
    report(f"antialias backward word planes mult={mult}")
    w = synth.small_workload()
    c = w.cards
    res = (w.H, w.W)
    shift = np.array([0.05, -0.03, 0.0, 0.0], np.float32)
    pos2 = np.concatenate([w.pos, w.pos + shift * w.pos[..., 3:4]], 0)
    cpos2 = np.concatenate([w.card_pos, w.card_pos + shift * w.card_pos[..., 3:4]], 0)
    K = 63 * mult + 2
    S = (K + 31) // 32
    target = oracle.dda_cluster_mask(res, pos2, w.edges, (w.edge2cluster * mult + 1).astype(np.int32), K)
    cards = oracle.triangle_cluster_mask(res, cpos2, c.tri, (c.tri2cluster * mult + 1).astype(np.int32), K)
    e2c = (c.edge2cluster * mult + 1).astype(np.int32)
    prim, normals = oracle.mark_discontinuity_edges(res, cpos2, c.tri, c.edges, c.edge2tri)
    fp, fa = oracle.dda_compute_fragments(res, cpos2, c.edges, prim)
    B, V, _ = cpos2.shape
    E = c.edges.shape[0]
    args = (fp.shape[0], T(fp, dev), T(fa, dev), res[0], res[1], S, T(cards, dev), T(target, dev), V,
            T(cpos2, dev), E, T(c.edges, dev), T(normals, dev), T(e2c, dev))
    for rho in (1.0, 0.5):
        want, ab = oracle.antialiased_cluster_bitset_grad(cards, target, cpos2, c.tri, c.edges,
                                                          c.edge2tri, e2c, 7.0, rho, True)
        assert np.abs(want[1]).max() > 0  # the second view contributes
        d_planes = torch.zeros((B, V, 4), device=dev)
        call("s2c_backward_antialiased_cluster_bitset_views", B, *args, d_planes, 7.0, rho)
        elementwise_close(d_planes.cpu().numpy(), want, ab, what=f"word planes S={S} rho={rho}")
        d_records = torch.zeros((B, V, 4), device=dev)
        call("s2c_backward_antialiased_cluster_bitset", *args, d_records, 7.0, rho)
        elementwise_close(d_records.cpu().numpy(), want, ab, what=f"records S={S} rho={rho}")


@pytest.mark.parametrize("mult,shuffle,persp", [(1, False, False), (1, True, False),
                                                (1, False, True), (17, False, False),
                                                (17, True, True)])
def test_antialias_marked_edges(dev, mult, shuffle, persp):
    """The gradient from the uncompacted marked edges (DDA walk inside the
    backward kernel, no fragment list) against the oracle's mark -> DDA ->
    backward composition: two views of the small workload, S = 2 (records)
    and 34 (word planes), card edges in file order (staged windows) and
    shuffled (waves mixing clusters), orthographic and perspective views
    (w in [0.5, 2], so the DDA divides and the depth test is live)."""
    report(f"antialias backward from marked edges mult={mult} shuffle={shuffle} persp={persp}")
    w = synth.small_workload()
    c = w.cards
    res = (w.H, w.W)
    shift = np.array([0.05, -0.03, 0.0, 0.0], np.float32)
    pos2 = np.concatenate([w.pos, w.pos + shift * w.pos[..., 3:4]], 0)
    cpos2 = np.concatenate([w.card_pos, w.card_pos + shift * w.card_pos[..., 3:4]], 0)
    if persp:
        ww = np.random.default_rng(8).uniform(0.5, 2.0, cpos2.shape[:2] + (1,)).astype(np.float32)
        cpos2 = (cpos2 * ww).astype(np.float32)
    edges, e2t, e2c = c.edges, c.edge2tri, c.edge2cluster
    if shuffle:
        perm = np.random.default_rng(4).permutation(edges.shape[0])
        edges, e2t, e2c = edges[perm].copy(), e2t[perm].copy(), e2c[perm].copy()
    K = 63 * mult + 2
    S = (K + 31) // 32
    target = oracle.dda_cluster_mask(res, pos2, w.edges, (w.edge2cluster * mult + 1).astype(np.int32), K)
    cards = oracle.triangle_cluster_mask(res, cpos2, c.tri, (c.tri2cluster * mult + 1).astype(np.int32), K)
    e2c = (e2c * mult + 1).astype(np.int32)
    B, V, _ = cpos2.shape
    E = edges.shape[0]
    tp, te = T(cpos2, dev), T(edges, dev)
    marked = torch.full((B * E,), -1, dtype=torch.int32, device=dev)
    normals = torch.zeros((B * E, 2), device=dev)
    call("s2c_mark_discontinuity_edges", res[0], res[1], B, V, tp, T(c.tri, dev), E, te,
         T(e2t, dev), marked, normals)
    prim, _ = oracle.mark_discontinuity_edges(res, cpos2, c.tri, edges, e2t)
    fp, fa = oracle.dda_compute_fragments(res, cpos2, edges, prim)
    for rho in (1.0, 0.5):
        want, ab = oracle.antialiased_cluster_bitset_grad(cards, target, cpos2, c.tri, edges, e2t,
                                                          e2c, 7.0, rho, True)
        assert np.abs(want[1]).max() > 0
        d_edges = torch.zeros((B, V, 4), device=dev)
        call("s2c_backward_antialiased_cluster_bitset_edges", B, res[0], res[1], S, T(cards, dev),
             T(target, dev), V, tp, E, te, marked, normals, T(e2c, dev), d_edges, 7.0, rho)
        elementwise_close(d_edges.cpu().numpy(), want, ab, what=f"marked edges S={S} rho={rho}")
        d_list = torch.zeros((B, V, 4), device=dev)
        call("s2c_backward_antialiased_cluster_bitset_views", B, fp.shape[0], T(fp, dev),
             T(fa, dev), res[0], res[1], S, T(cards, dev), T(target, dev), V, tp, E, te, normals,
             T(e2c, dev), d_list, 7.0, rho)
        # both forms carry summation-order rounding: twice the oracle bound
        elementwise_close(d_edges.cpu().numpy(), d_list.cpu().numpy(), ab, rtol=0.0, k=128,
                          what=f"marked edges vs list S={S} rho={rho}")


def test_antialias_marked_edges_full_workload(dev):
    """Both gradient forms at full size: 4 views at 1024^2, 1,024 clusters
    (S = 32, the word-plane path), 1,024 cards x 20 quads, masks rendered by
    the library's fused DDA and triangle masks (bit-exact with the oracle's,
    checked elsewhere).  Both forms are compared element by element with the
    oracle's gradient of the same masks, and with each other."""
    report("antialias backward from marked edges, full workload")
    w = synth.full_workload()
    c = w.cards
    res = (w.H, w.W)
    B, V, _ = w.card_pos.shape
    E = c.edges.shape[0]
    target = T(fused_dda_mask(dev, res, w.pos, w.edges, w.edge2cluster, 1024), dev)
    tp, te, tt = T(w.card_pos, dev), T(c.edges, dev), T(c.tri, dev)
    cards = torch.zeros_like(target)
    call("s2c_triangle_cluster_mask", B, res[0], res[1], V, tp, c.tri.shape[0], tt,
         T(c.tri2cluster, dev), target.shape[3], cards)
    marked = torch.full((B * E,), -1, dtype=torch.int32, device=dev)
    normals = torch.zeros((B * E, 2), device=dev)
    call("s2c_mark_discontinuity_edges", res[0], res[1], B, V, tp, tt, E, te, T(c.edge2tri, dev),
         marked, normals)
    prim = torch.empty_like(marked)
    m = call_count("s2c_compact_valid_ints", B * E, marked, prim)
    _, _, _, _, fp, fa = dda_gpu(dev, res, w.card_pos, c.edges, prim[:m].cpu().numpy())
    S = target.shape[3]
    e2c = T(c.edge2cluster, dev)
    d_list = torch.zeros((B, V, 4), device=dev)
    call("s2c_backward_antialiased_cluster_bitset_views", B, fp.shape[0], T(fp, dev), T(fa, dev),
         res[0], res[1], S, cards, target, V, tp, E, te, normals, e2c, d_list, 7.0, 0.5)
    d_edges = torch.zeros((B, V, 4), device=dev)
    call("s2c_backward_antialiased_cluster_bitset_edges", B, res[0], res[1], S, cards, target, V,
         tp, E, te, marked, normals, e2c, d_edges, 7.0, 0.5)
    got_list, got_edges = d_list.cpu().numpy(), d_edges.cpu().numpy()
    assert np.abs(got_list).max() > 0 and fp.shape[0] > 100000
    want, ab = oracle.antialiased_cluster_bitset_grad(cards.cpu().numpy(), target.cpu().numpy(),
                                                      w.card_pos, c.tri, c.edges, c.edge2tri,
                                                      c.edge2cluster, 7.0, 0.5, True)
    elementwise_close(got_list, want, ab, what="full workload, list form vs oracle")
    elementwise_close(got_edges, want, ab, what="full workload, marked edges vs oracle")
    elementwise_close(got_edges, got_list, ab, rtol=0.0, k=128,
                      what="full workload, marked edges vs list form")


def test_antialias_marked_edges_empty(dev):
    """No marked edge, no edges, no views: d_pos is left untouched."""
    report("antialias backward from marked edges: empty inputs")
    H, W, S, V, E = 16, 16, 2, 4, 3
    bits = torch.zeros((1, H, W, S), dtype=torch.int32, device=dev)
    pos = torch.zeros((1, V, 4), device=dev)
    edges = torch.tensor([[0, 1], [1, 2], [2, 3]], dtype=torch.int32, device=dev)
    marked = torch.full((E,), -1, dtype=torch.int32, device=dev)
    normals = torch.zeros((E, 2), device=dev)
    e2c = torch.zeros(E, dtype=torch.int32, device=dev)
    d = torch.full((1, V, 4), 3.0, device=dev)
    call("s2c_backward_antialiased_cluster_bitset_edges", 1, H, W, S, bits, bits, V, pos, E, edges,
         marked, normals, e2c, d, 7.0, 1.0)
    call("s2c_backward_antialiased_cluster_bitset_edges", 1, H, W, S, bits, bits, V, pos, 0, edges,
         marked, normals, e2c, d, 7.0, 1.0)
    call("s2c_backward_antialiased_cluster_bitset_edges", 0, H, W, S, bits, bits, V, pos, E, edges,
         marked, normals, e2c, d, 7.0, 1.0)
    torch.cuda.synchronize()
    assert (d == 3.0).all()


# ---------------------------------------------------------------- curves


def test_smooth_and_skinning(dev):
    report("smooth_strands / skinning")
    rng = np.random.default_rng(3)
    x = synth.generate_strands(300, 40, 0.004, 3)
    s = oracle.compute_arclength(x)
    for beta in (0.002, 0.01, 0.05):
        out = torch.zeros(x.shape, device=dev)
        call("s2c_smooth_strands", 300, 40, T(x, dev), T(s, dev), out, beta)
        want, ab = oracle.smooth_strands(x, s, beta, True)
        elementwise_close(out.cpu().numpy(), want, ab, k=8, what=f"smooth_strands beta={beta}")
    h = x + rng.normal(scale=1e-3, size=x.shape).astype(np.float32)
    cn = x + rng.normal(scale=1e-3, size=x.shape).astype(np.float32)
    q, _ = np.linalg.qr(rng.normal(size=(300, 40, 3, 3)))
    R = q.astype(np.float32)
    out = torch.zeros(x.shape, device=dev)
    call("s2c_skinning", 300, 40, T(x, dev), T(s, dev), T(h, dev), T(cn, dev), T(R, dev), out, 0.01)
    want, ab = oracle.skinning(x, s, h, cn, R, 0.01, True)
    elementwise_close(out.cpu().numpy(), want, ab, k=8, what="skinning")


def test_smooth_and_skinning_edge_cases(dev):
    """smooth_strands / skinning (curve_utils.cu:135-166) where the weights
    are special: strands of 1 and 2 points; repeated points (zero arc-length
    steps, so weight 1 between distinct points); beta = 10, where every
    point of a strand weighs about 1; beta = 1e-4, where each point keeps
    little but its repeats; and beta = 0, where the weights are
    exp(-(0 * inf)) = NaN and the weight-sum test fails, so smoothing leaves
    the zero-filled output as it is and skinning keeps each point; beside a
    4,096 x 100 batch."""
    report("smooth_strands / skinning edge cases")
    rng = np.random.default_rng(29)
    for N, P in ((5, 1), (7, 2), (64, 40), (4096, 100)):
        x = np.cumsum(rng.normal(scale=0.004, size=(N, P, 3)), 1).astype(np.float32)
        if P == 40:
            x[:, 10:15] = x[:, 10:11]  # five equal points: zero arc-length steps
            x[:, 30:32] = x[:, 29:30]
        s = oracle.compute_arclength(x)
        h = x + rng.normal(scale=1e-3, size=x.shape).astype(np.float32)
        cn = x + rng.normal(scale=1e-3, size=x.shape).astype(np.float32)
        q, _ = np.linalg.qr(rng.normal(size=(N, P, 3, 3)))
        R = np.ascontiguousarray(q.astype(np.float32))
        for beta in ((0.0, 1e-4, 0.01, 10.0) if N < 4096 else (0.01,)):
            what = f"N={N} P={P} beta={beta}"
            out = torch.zeros(x.shape, device=dev)
            call("s2c_smooth_strands", N, P, T(x, dev), T(s, dev), out, beta)
            want, ab = oracle.smooth_strands(x, s, beta, True)
            if beta == 0.0:
                assert not want.any()
                np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=what)
            else:
                elementwise_close(out.cpu().numpy(), want, ab, k=8, what="smooth " + what)
            out = torch.zeros(x.shape, device=dev)
            call("s2c_skinning", N, P, T(x, dev), T(s, dev), T(h, dev), T(cn, dev), T(R, dev),
                 out, beta)
            want, ab = oracle.skinning(x, s, h, cn, R, beta, True)
            if beta == 0.0:  # skinning keeps the point itself where the sum test fails
                np.testing.assert_array_equal(want, x)
                np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=what)
            else:
                elementwise_close(out.cpu().numpy(), want, ab, k=8, what="skinning " + what)
