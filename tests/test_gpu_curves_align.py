"""GPU tests of the Bishop-frame kernel and the alignment helpers built on it
(curve_utils.py:104-165, 242-372, 374-644)."""
import numpy as np
import pytest
import torch

import oracle
from conftest import report
from strns2cards_amd import curve_utils as cu
from strns2cards_amd import synth

pytestmark = pytest.mark.gpu


def test_bishop_kernel_matches_oracle(dev):
    report("curves: Bishop frame kernel vs oracle (2000 strands x 32)")
    x = synth.generate_strands(2000, 32, 0.004, 0).astype(np.float32)
    T, N, B = cu.compute_bishop_frame(torch.from_numpy(x).to(dev))
    To, No, Bo = oracle.bishop_frame(x)
    for got, want in ((T, To), (N, No), (B, Bo)):
        g = got.cpu().numpy()
        assert np.abs(g - want).max() < 1e-5
    # degenerate inputs: one vertex, repeated vertices
    y = np.zeros((3, 5, 3), np.float32)
    y[1, :, 0] = np.arange(5)
    y[2, 2:, 1] = 1.0
    for got, want in zip(cu.compute_bishop_frame(torch.from_numpy(y).to(dev)),
                         oracle.bishop_frame(y)):
        np.testing.assert_allclose(got.cpu().numpy(), want, atol=1e-6)
    one = cu.compute_bishop_frame(torch.ones((2, 1, 3), device=dev))
    np.testing.assert_allclose(one[1].cpu().numpy()[:, 0], [[1, 0, 0]] * 2)


def test_bishop_host_input_round_trip(dev):
    x = synth.generate_strands(10, 16, 0.004, 2).astype(np.float64)
    T, N, B = cu.compute_bishop_frame(torch.from_numpy(x))  # float64 on the host
    assert T.dtype == torch.float64 and not T.is_cuda
    To, _, _ = oracle.bishop_frame(x.astype(np.float32))
    assert np.abs(T.numpy() - To).max() < 1e-5


def test_alignment_rotation_maps_frames(dev):
    report("curves: alignment rotations")
    x = torch.from_numpy(synth.generate_strands(300, 24, 0.004, 1).astype(np.float32)).to(dev)
    Td = torch.tensor([[0.0, 1.0, 0.0]], device=dev)
    R = cu.compute_alignment_rotation(x, Td).double()
    T, N, B = (f.double() for f in cu.compute_bishop_frame(x))
    assert torch.allclose((R @ T[..., None])[..., 0], Td.double().expand_as(T), atol=1e-5)
    nroot = (R[:, :1] @ N[:, :1, :, None])[..., 0]
    assert torch.allclose((R @ N[..., None])[..., 0], nroot.expand_as(N), atol=1e-5)
    eye = torch.eye(3, dtype=torch.float64, device=dev).expand_as(R)
    assert torch.allclose(R @ R.transpose(-1, -2), eye, atol=1e-5)


def test_aligned_strands_closed_form(dev):
    x = torch.from_numpy(synth.generate_strands(50, 20, 0.004, 3).astype(np.float32)).to(dev)
    roots = torch.randn(50, 3, device=dev)
    Td = torch.tensor([[0.0, 0.0, 1.0]], device=dev)
    a = cu.compute_aligned_strands(x, roots, Td)
    s = torch.cat([torch.zeros(50, 1, device=dev),
                   torch.cumsum((x[:, 1:] - x[:, :-1]).norm(dim=-1), 1)], 1)
    assert torch.allclose(a[..., 2], roots[:, None, 2] + s, atol=1e-6)
    assert torch.allclose(a[..., :2], roots[:, None, :2].expand(50, 20, 2))


def _straight_wisp(rng, n, v, d, spread=0.05):
    roots = np.zeros((n, 3))
    roots[:, 0] = rng.uniform(-spread, spread, n)
    roots[:, 1] = rng.uniform(-spread, spread, n)
    step = rng.uniform(0.003, 0.006, n)
    return (roots[:, None, :] + (np.arange(v)[None, :, None] * step[:, None, None]) * d).astype(
        np.float32), step


def test_align_wisps_straightens_straight_wisps(dev):
    """Straight strands tilted out of their roots' plane align to straight
    vertical lines with their segment lengths kept (the skinning of a
    straight handle is a rigid motion onto the canonical line)."""
    report("curves: align_wisps on straight wisps")
    rng = np.random.default_rng(0)
    d = np.array([0.0, 1.0, 0.1]) / np.linalg.norm([0.0, 1.0, 0.1])
    wisps, steps = [], []
    for k in range(5):
        w, st = _straight_wisp(rng, 6 + 3 * k, 24, d)
        wisps.append(w + rng.normal(size=3).astype(np.float32))
        steps.append(st)
    aligned, handles0, longest = cu.align_wisps(wisps, beta=0.01)
    steps = np.concatenate(steps)
    a = aligned.cpu().numpy()
    seg = np.diff(a, axis=1)
    want = steps[:, None, None] * np.array([0.0, 1.0, 0.0])[None, None]
    assert np.abs(seg - want).max() < 2e-5
    # handles are the unrotated smoothing of the input; longest = first maximal strand per wisp
    allx = np.concatenate(wisps)
    assert handles0.shape == allx.shape
    offs = np.concatenate([[0], np.cumsum([w.shape[0] for w in wisps])[:-1]])
    lens = np.linalg.norm(np.diff(allx, axis=1), axis=2).sum(1)
    for k, w in enumerate(wisps):
        assert int(longest[k]) == offs[k] + int(np.argmax(lens[offs[k]:offs[k] + w.shape[0]]))


def _varied(x, rng):
    """Strands of distinct lengths (so the longest one is well defined)."""
    k = rng.uniform(0.6, 1.4, size=(x.shape[0], 1, 1))
    return (x[:, :1] + k * (x - x[:, :1])).astype(np.float32)


def _rigid(rng, x):
    q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return (x @ q.T + rng.normal(size=3)).astype(np.float32)


def test_align_wisp_rigid_invariance(dev):
    """align_wisp only sees a wisp's shape: the centring, the roots'
    principal axis and the guide frame all move with a rigid motion of the
    input, so a rotated and translated copy aligns to the same strands.
    (Perfectly straight guides are excluded: their first two tangents are
    parallel and the root binormal is rounding noise, in the reference too.)"""
    report("curves: align_wisp rigid-motion invariance")
    rng = np.random.default_rng(1)
    x = _varied(synth.generate_strands(40, 32, 0.004, 5), rng)
    a1, h1, l1 = cu.align_wisp(torch.from_numpy(x).to(dev), beta=0.01)
    a2, h2, l2 = cu.align_wisp(torch.from_numpy(_rigid(rng, x)).to(dev), beta=0.01)
    assert int(l1) == int(l2)
    assert np.abs(a1.cpu().numpy() - a2.cpu().numpy()).max() < 5e-5
    assert np.abs(h1.cpu().numpy() - h2.cpu().numpy()).max() < 5e-5
    # the rotated handles' root tangent of the guide is the canonical one
    g = h1[int(l1)].cpu().numpy()
    assert np.isfinite(a1.cpu().numpy()).all()
    t = g[1] - g[0]
    assert t[1] < 0  # points along -y


def test_align_wisps_rigid_invariance(dev):
    report("curves: align_wisps rigid-motion invariance per wisp")
    rng = np.random.default_rng(2)
    wisps = [_varied(synth.generate_strands(n, 24, 0.004, 10 + n), rng) for n in (2, 5, 17, 30)]
    moved = [_rigid(rng, w) for w in wisps]
    a1, _, l1 = cu.align_wisps(wisps, beta=0.01)
    a2, _, l2 = cu.align_wisps(moved, beta=0.01)
    assert torch.equal(l1.cpu(), l2.cpu())
    # wisps of >= 3 strands are invariant; smaller ones use fixed axes (not invariant)
    sizes = np.cumsum([0] + [w.shape[0] for w in wisps])
    d = np.abs(a1.cpu().numpy() - a2.cpu().numpy())
    assert d[sizes[1]:].max() < 5e-5


def test_strands_to_tube_layout(dev):
    x = synth.generate_strands(3, 6, 0.004, 4)
    attr = np.arange(3 * 6 * 2, dtype=np.float32).reshape(3, 6, 2)
    verts, faces, ex = cu.strands_to_tube(x, radius=0.01, n_circle=5, attr=attr)
    assert verts.shape == (3 * 6 * 5, 3) and faces.dtype == np.int32
    assert faces.shape == (3 * 2 * 5 * 5, 3)
    ring = verts.reshape(3, 6, 5, 3) - x[:, :, None, :]
    np.testing.assert_allclose(np.linalg.norm(ring, axis=-1), 0.01, rtol=1e-5)
    np.testing.assert_array_equal(ex.reshape(3, 6, 5, 2), np.repeat(attr[:, :, None], 5, axis=2))
    # template of one strand with S = 2, n_circle = 3
    _, f = cu.strands_to_tube(x[:1, :2], n_circle=3)
    np.testing.assert_array_equal(f, [[0, 3, 4], [1, 4, 5], [2, 5, 3],
                                      [0, 4, 1], [1, 5, 2], [2, 3, 0]])


def test_statistical_outlier_mask(dev):
    rng = np.random.default_rng(2)
    p = rng.normal(size=(500, 3)).astype(np.float32) * 0.01
    p[:5] += 10.0 * rng.normal(size=(5, 3)).astype(np.float32)
    m = cu.statistical_outlier_mask(p, n_neighbors=5, ratio=2.0)
    assert not m[:5].any() and m[5:].mean() > 0.9
    # brute force in float64
    d = np.linalg.norm(p[:, None].astype(np.float64) - p[None].astype(np.float64), axis=-1)
    near = np.sort(d, axis=1)[:, 1:6].mean(1)
    np.testing.assert_array_equal(m, near < 2.0 * near.mean())
    assert cu.statistical_outlier_mask(p[:1]).tolist() == [True]
