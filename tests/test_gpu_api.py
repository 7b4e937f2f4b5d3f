"""GPU tests of the Python API (strns2cards_amd, mirroring pystrns2cards)
against the CPU oracle, including autograd wrappers and error behaviour."""
import numpy as np
import pytest
import torch

import oracle
import strns2cards_amd as s2c
from conftest import elementwise_close, report
from strns2cards_amd import synth

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def small():
    return synth.small_workload()


def test_dda_api(dev, small):
    report("api: dda_compute_fragments / cluster_mask_from_fragments / dda_cluster_mask")
    w = small
    res = (w.H, w.W)
    fp, fa = s2c.dda_compute_fragments(res, T(w.pos, dev), T(w.edges, dev))
    o_fp, o_fa = oracle.dda_compute_fragments(res, w.pos, w.edges)
    np.testing.assert_array_equal(fp.cpu().numpy(), o_fp)
    np.testing.assert_array_equal(fa.cpu().numpy(), o_fa)
    bits = s2c.cluster_mask_from_fragments(res, fp, fa, T(w.edge2cluster, dev), cluster_count=64)
    want = oracle.dda_cluster_mask(res, w.pos, w.edges, w.edge2cluster, 64)
    np.testing.assert_array_equal(bits.cpu().numpy(), want)
    fused = s2c.dda_cluster_mask(res, T(w.pos, dev), T(w.edges, dev), T(w.edge2cluster, dev), 64)
    np.testing.assert_array_equal(fused.cpu().numpy(), want)


def test_dda_interpolate_autograd(dev, small):
    report("api: DDAInterpolateAttributes fwd/bwd")
    w = small
    res = (w.H, w.W)
    fp, fa = oracle.dda_compute_fragments(res, w.pos, w.edges)
    rast = oracle.depth_test(1, res, fp, fa)
    rng = np.random.default_rng(0)
    attr = rng.normal(size=(w.pos.shape[1], 3)).astype(np.float32)
    ta = T(attr, dev).requires_grad_(True)
    img = s2c.dda_interpolate_attributes(T(rast, dev), ta, T(w.edges, dev))
    np.testing.assert_array_equal(img.detach().cpu().numpy(),
                                  oracle.dda_interpolate_attributes(rast, attr, w.edges))
    g = rng.normal(size=img.shape).astype(np.float32)
    img.backward(T(g, dev))
    want, ab = oracle.backward_dda_interpolate_attributes(rast, g, w.edges, attr.shape[0], True)
    elementwise_close(ta.grad.cpu().numpy(), want, ab, what="DDA interpolation backward")


def test_rasterize_and_interpolate(dev, small):
    report("api: compute_fragments / depth_test / rasterize / interpolate")
    w = small
    res = (w.H, w.W)
    c = w.cards
    pos, tri = T(w.card_pos, dev), T(c.tri, dev)
    fp, fa = s2c.compute_fragments(res, pos, tri)
    o_fp, o_fa = oracle.compute_fragments(res, w.card_pos, c.tri)
    np.testing.assert_array_equal(fp.cpu().numpy(), o_fp)
    np.testing.assert_array_equal(fa.cpu().numpy(), o_fa)
    rast = s2c.rasterize(res, pos, tri)
    o_rast = oracle.rasterize(res, w.card_pos, c.tri)
    np.testing.assert_array_equal(rast.cpu().numpy(), o_rast)
    uv = np.random.default_rng(1).uniform(size=(c.verts.shape[0], 2)).astype(np.float32)
    tu = T(uv, dev).requires_grad_(True)
    img = s2c.interpolate_triangle_attributes(rast, tu, tri)
    np.testing.assert_array_equal(img.detach().cpu().numpy(),
                                  oracle.interpolate_triangle_attributes(o_rast, uv, c.tri))
    img.sum().backward()
    want, ab = oracle.backward_interpolate_triangle_attributes(
        o_rast, np.ones(img.shape, np.float32), c.tri, uv.shape[0], True)
    elementwise_close(tu.grad.cpu().numpy(), want, ab, what="triangle interpolation backward")
    bits = s2c.cluster_mask_from_fragments(res, fp, fa, T(c.tri2cluster, dev), 64)
    fused = s2c.triangle_cluster_mask(res, pos, tri, T(c.tri2cluster, dev), 64)
    want_bits = oracle.triangle_cluster_mask(res, w.card_pos, c.tri, c.tri2cluster, 64)
    np.testing.assert_array_equal(bits.cpu().numpy(), want_bits)
    np.testing.assert_array_equal(fused.cpu().numpy(), want_bits)


def test_dda_cluster_mask_two_streams(dev):
    """The fused DDA mask keeps its binning buffers per (device, stream):
    two streams working at the same time, on different workloads and
    resolutions (so the second call on each stream grows its buffers), give
    the oracle's masks."""
    report("dda_cluster_mask on two streams at once")
    rng = np.random.default_rng(77)
    cases = []
    for res, n, seg in (((300, 260), 3000, 20.0), ((512, 512), 12000, 40.0)):
        H, W = res
        p0 = rng.uniform(0, [W, H], size=(n, 2))
        p1 = p0 + rng.normal(scale=seg, size=(n, 2))
        pts = np.stack([p0, p1], 1).reshape(-1, 2)
        pos = np.zeros((2, 2 * n, 4), np.float32)
        pos[:, :, 0] = 2 * pts[:, 0] / W - 1
        pos[:, :, 1] = 1 - 2 * pts[:, 1] / H
        pos[1, :, 0] += 0.01
        pos[..., 3] = 1
        edges = np.arange(2 * n, dtype=np.int32).reshape(n, 2)
        e2c = rng.integers(0, 200, size=n).astype(np.int32)
        cases.append((res, pos, edges, e2c, oracle.dda_cluster_mask(res, pos, edges, e2c, 200)))
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    for order in ((0, 1), (1, 0)):  # each stream sees the small and the large case
        outs = []
        for si, ci in enumerate(order):
            res, pos, edges, e2c, _ = cases[ci]
            with torch.cuda.stream(streams[si]):
                outs.append(s2c.dda_cluster_mask(res, T(pos, dev), T(edges, dev), T(e2c, dev),
                                                 cluster_count=200))
        torch.cuda.synchronize()
        for out, ci in zip(outs, order):
            np.testing.assert_array_equal(out.cpu().numpy(), cases[ci][4])
    # releasing a stream's buffers: the next call on it allocates them again
    from capi import call
    for st in streams:
        call("s2c_dda_cluster_mask_release", stream=st.cuda_stream)
    res, pos, edges, e2c, want = cases[1]
    with torch.cuda.stream(streams[0]):
        out = s2c.dda_cluster_mask(res, T(pos, dev), T(edges, dev), T(e2c, dev), cluster_count=200)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_bitset_api(dev, small):
    report("api: accumulate_bitset_rgb / popcount_bitset")
    w = small
    bits = oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, 64)
    rgb = np.random.default_rng(2).uniform(size=(64, 3)).astype(np.float32)
    out = s2c.accumulate_bitset_rgb(T(bits, dev), T(rgb, dev))
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.accumulate_bitset_rgb(bits, rgb))
    cnt = s2c.popcount_bitset(T(bits, dev))
    np.testing.assert_array_equal(cnt.cpu().numpy(), oracle.popcount_bitset(bits))


def test_antialias_loss_api(dev, small):
    report("api: mark_discontinuity_edges / antialiased_cluster_bitset_loss")
    w = small
    c = w.cards
    res = (w.H, w.W)
    target = oracle.dda_cluster_mask(res, w.pos, w.edges, w.edge2cluster, 64)
    bits = oracle.triangle_cluster_mask(res, w.card_pos, c.tri, c.tri2cluster, 64)
    pos = T(w.card_pos, dev)
    prim, normals = s2c.mark_discontinuity_edges(res, pos, T(c.tri, dev), T(c.edges, dev),
                                                 T(c.edge2tri, dev))
    o_prim, o_normals = oracle.mark_discontinuity_edges(res, w.card_pos, c.tri, c.edges, c.edge2tri)
    np.testing.assert_array_equal(prim.cpu().numpy(), o_prim)
    np.testing.assert_array_equal(normals.cpu().numpy(), o_normals)
    pos.requires_grad_(True)
    loss = s2c.antialiased_cluster_bitset_loss(T(bits, dev), T(target, dev), pos, T(c.tri, dev),
                                               T(c.edges, dev), T(c.edge2tri, dev),
                                               T(c.edge2cluster, dev), 7.0, 0.5, True)
    (3.0 * loss).backward()
    want, ab = oracle.antialiased_cluster_bitset_grad(bits, target, w.card_pos, c.tri, c.edges,
                                                      c.edge2tri, c.edge2cluster, 7.0, 0.5, True)
    want, ab = 3.0 * want, 3.0 * ab
    elementwise_close(pos.grad.cpu().numpy(), want, ab, what="antialiased loss gradient")
    assert loss.item() == oracle.popcount_bitset(np.bitwise_xor(bits, target)).sum()
    pos.grad = None
    loss = s2c.antialiased_cluster_bitset_loss(T(bits, dev), T(target, dev), pos, T(c.tri, dev),
                                               T(c.edges, dev), T(c.edge2tri, dev),
                                               T(c.edge2cluster, dev), 7.0, 0.5, True, fused=True)
    (3.0 * loss).backward()
    elementwise_close(pos.grad.cpu().numpy(), want, ab, what="antialiased loss gradient, fused")
    assert loss.item() == oracle.popcount_bitset(np.bitwise_xor(bits, target)).sum()


def test_curve_api(dev):
    report("api: smooth_strands / skinning")
    x = synth.generate_strands(200, 30, 0.004, 9)
    s = oracle.compute_arclength(x)
    tx = T(x, dev)
    np.testing.assert_allclose(s2c.compute_arclength(tx).cpu().numpy(), s, rtol=1e-5, atol=1e-7)
    got = s2c.smooth_strands(tx, 0.01).cpu().numpy()
    s_dev = s2c.compute_arclength(tx).cpu().numpy()
    want, ab = oracle.smooth_strands(x, s_dev, 0.01, True)
    elementwise_close(got, want, ab, k=8, what="smooth_strands")
    R = np.broadcast_to(np.eye(3, dtype=np.float32), (200, 30, 3, 3)).copy()
    got = s2c.skinning(tx, tx, tx, T(R, dev), 0.01).cpu().numpy()
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-6)  # R = I, h = c -> identity


def test_api_errors(dev):
    report("api: argument errors")
    pos = torch.zeros((1, 4, 4), device=dev)
    edges = torch.zeros((2, 2), dtype=torch.int32, device=dev)
    with pytest.raises(ValueError):
        s2c.dda_compute_fragments((8, 8), pos.cpu(), edges)
    with pytest.raises(ValueError):
        s2c.dda_compute_fragments((8, 8), pos.double(), edges)
    with pytest.raises(ValueError):
        s2c.compute_fragments((8, 8), pos, edges)
    with pytest.raises(TypeError):
        s2c.popcount_bitset(torch.zeros((1, 2, 2, 1), dtype=torch.int64, device=dev))
    with pytest.raises(RuntimeError):  # zero fragments: .max() of an empty tensor
        s2c.cluster_mask_from_fragments((8, 8), torch.zeros((0, 3), dtype=torch.int32, device=dev),
                                        torch.zeros((0, 4), device=dev),
                                        torch.zeros(1, dtype=torch.int32, device=dev))
    with pytest.raises(RuntimeError):
        from strns2cards_amd import _core
        _core.depth_test(8, 8, torch.zeros((3, 3), dtype=torch.int32, device=dev),
                         torch.zeros((2, 4), device=dev), torch.zeros((1, 8, 8, 4), device=dev))
