"""GPU tests of the hand-written coverage adjoint (``CardCoverage``).

* forward and backward against the float64 oracle (normwise 1e-5);
* the adjoint against central finite differences of the float64 loss on
  64 random card vertices, at h = 1e-4 (truncation-limited: the coverage is
  only C1 where a vertex or an edge crosses a pixel line, which bounds how
  well a 1e-4 step can agree) and at h = 1e-6 (where the difference
  quotient converges and the bar is the fp32 adjoint's own accuracy);
* the fitting loop over the parameters (centre-lines and widths).
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import report
from strns2cards_amd import adjoint, synth

pytestmark = pytest.mark.gpu


def normwise(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.fixture(scope="module")
def small():
    return synth.small_workload()


def _clip64(verts, mvp):
    vh = np.concatenate([verts, np.ones((verts.shape[0], 1))], 1)
    return np.einsum("bij,vj->bvi", mvp, vh)


def _random_scene(rng, B, H, W, T):
    """Random triangles with perspective w, some off screen, some behind the
    eye and some degenerate."""
    V = 3 * T
    pos = np.zeros((B, V, 4), np.float32)
    w = rng.uniform(0.5, 2.0, size=(B, V))
    pos[..., 0] = rng.uniform(-1.3, 1.3, size=(B, V)) * w
    pos[..., 1] = rng.uniform(-1.3, 1.3, size=(B, V)) * w
    pos[..., 2] = rng.uniform(-0.5, 0.5, size=(B, V)) * w
    pos[..., 3] = w
    # shrink most triangles around their first vertex so they stay small
    for b in range(B):
        for t in range(T):
            s = rng.uniform(0.02, 0.25)
            c = pos[b, 3 * t].copy()
            for k in (1, 2):
                pos[b, 3 * t + k, :3] = c[:3] + s * (pos[b, 3 * t + k, :3] - c[:3])
    pos[0, 0, 3] = -1.0                # behind the eye
    pos[0, 4] = pos[0, 3]              # degenerate (two equal vertices)
    tri = np.arange(V, dtype=np.int32).reshape(T, 3)
    return pos, tri


def test_forward_parity_random(dev):
    report("adjoint: coverage forward vs oracle (random perspective triangles)")
    rng = np.random.default_rng(0)
    B, H, W, T = 2, 96, 160, 300
    pos, tri = _random_scene(rng, B, H, W, T)
    got = adjoint.card_coverage(torch.from_numpy(pos).to(dev), torch.from_numpy(tri).to(dev),
                                (H, W)).cpu().numpy()
    want = oracle.card_coverage((H, W), pos.astype(np.float64), tri)
    assert normwise(got, want) < 1e-5
    # fp32 screen coordinates carry ~1e-5 px of rounding at this size
    assert np.abs(got - want).max() < 5e-5 * max(1.0, np.abs(want).max())


def test_forward_parity_cards(dev, small):
    report("adjoint: coverage forward vs oracle (64 cards x 16 quads, 256^2)")
    H, W = small.H, small.W
    pos = torch.from_numpy(small.card_pos).to(dev)
    tri = torch.from_numpy(small.cards.tri).to(dev)
    got = adjoint.card_coverage(pos, tri, (H, W)).cpu().numpy()
    want = oracle.card_coverage((H, W), small.card_pos.astype(np.float64), small.cards.tri)
    assert normwise(got, want) < 1e-5
    assert np.abs(got - want).max() < 5e-5 * max(1.0, np.abs(want).max())
    # sum of coverage = total card area on screen (all cards are inside)
    assert abs(got.sum() - want.sum()) < 1e-5 * want.sum()


def test_backward_parity(dev, small):
    report("adjoint: coverage backward vs oracle gradient")
    H, W = small.H, small.W
    rng = np.random.default_rng(1)
    target = rng.uniform(0.0, 1.0, size=(1, H, W))
    cov = oracle.card_coverage((H, W), small.card_pos.astype(np.float64), small.cards.tri)
    d_cov = 2.0 * (cov - target)
    want = oracle.card_coverage_grad((H, W), small.card_pos.astype(np.float64), small.cards.tri,
                                     d_cov)
    pos = torch.from_numpy(small.card_pos).to(dev)
    d_pos = torch.zeros_like(pos)
    from strns2cards_amd import _core
    _core.card_coverage_backward(pos, torch.from_numpy(small.cards.tri).to(dev),
                                 torch.from_numpy(d_cov.astype(np.float32)).to(dev), d_pos)
    got = d_pos.cpu().numpy()
    assert normwise(got, want) < 1e-5
    assert np.all(got[..., 2] == 0)


def test_backward_parity_random(dev):
    report("adjoint: coverage backward vs oracle (random perspective triangles)")
    rng = np.random.default_rng(2)
    B, H, W, T = 2, 96, 160, 300
    pos, tri = _random_scene(rng, B, H, W, T)
    p = torch.from_numpy(pos).to(dev).requires_grad_(True)
    tgt = torch.from_numpy(rng.uniform(0, 1, size=(B, H, W)).astype(np.float32)).to(dev)
    cov = adjoint.card_coverage(p, torch.from_numpy(tri).to(dev), (H, W))
    ((cov - tgt) ** 2).sum().backward()
    d_cov = 2.0 * (cov.detach().cpu().numpy().astype(np.float64) - tgt.cpu().numpy())
    want = oracle.card_coverage_grad((H, W), pos.astype(np.float64), tri, d_cov)
    assert normwise(p.grad.cpu().numpy(), want) < 1e-5


def _fd_check(dev, wl, views, n_vert, seed):
    """Adjoint gradient w.r.t. world-space card vertices vs central
    differences of the float64 loss sum (C - D)^2 / (B H W), taken at the
    fp32 clip-space positions the GPU used.  Returns, per step h, the
    normwise error of the GPU adjoint and of the float64 oracle adjoint."""
    H, W = wl.H, wl.W
    cards = wl.cards
    mvp = wl.mvp[:views]
    rng = np.random.default_rng(seed)
    target = torch.from_numpy(rng.uniform(0, 1, size=(views, H, W)).astype(np.float32)).to(dev)
    fit = adjoint.CardDensityFit.from_cards(cards, mvp, target)
    verts = fit.verts().detach().requires_grad_(True)
    loss = fit.loss(verts)
    loss.backward()
    ad = verts.grad.cpu().numpy().astype(np.float64)
    mvp64 = fit.mvpT.transpose(1, 2).cpu().numpy().astype(np.float64)
    with torch.no_grad():
        clip = fit.clip(verts)
        cov = adjoint.card_coverage(clip, fit.tri, (H, W)).cpu().numpy()
    pos64 = clip.cpu().numpy().astype(np.float64)
    tri = cards.tri
    tgt = target.cpu().numpy()
    n = views * H * W
    vs = rng.choice(pos64.shape[1], n_vert, replace=False)
    exact = oracle.card_coverage_grad((H, W), pos64, tri, 2.0 * (cov - tgt) / n)
    exact = np.einsum("bvi,bij->vj", exact, mvp64)[vs, :3]
    errs = {"gpu_vs_f64": normwise(ad[vs], exact)}
    for h in (1e-4, 1e-6):
        fd = np.zeros((n_vert, 3))
        for i, v in enumerate(vs):
            for k in range(3):
                d = np.zeros_like(pos64)
                d[:, v] = h * mvp64[:, :, k]
                fd[i, k] = oracle.coverage_loss_difference((H, W), pos64, tri, tgt, cov, d) / (h * n)
        errs[h] = (normwise(ad[vs], fd), normwise(exact, fd))
    return errs


def _check_fd(errs):
    (g4, e4), (g6, e6) = errs[1e-4], errs[1e-6]
    report(f"  GPU vs float64 adjoint {errs['gpu_vs_f64']:.3g}; normwise |ad - fd| / |fd|: "
           f"h=1e-4 gpu {g4:.3g} (float64 adjoint {e4:.3g}), "
           f"h=1e-6 gpu {g6:.3g} (float64 adjoint {e6:.3g})")
    # the GPU adjoint against the float64 adjoint on the same vertices
    assert errs["gpu_vs_f64"] < 1e-5
    # Central differences converge to the adjoint.  They are limited by
    # truncation at kinks: the coverage is only C1 where a vertex or an edge
    # crosses a pixel line or corner, and a step of h moves every edge of
    # the vertex by h * (pixels per unit) across the pixel grid.
    assert e4 < 2e-2 and g4 < 2e-2
    assert e6 < 1e-4 and g6 < 1e-4
    assert e6 < 0.1 * e4


def test_finite_differences_small(dev, small):
    report("adjoint: central differences on 64 card vertices (64 cards, 256^2)")
    _check_fd(_fd_check(dev, small, 1, 64, 3))


def test_fit_reduces_loss(dev, small):
    report("adjoint: 30 Adam steps on centre-lines and widths (small workload)")
    pos = torch.from_numpy(small.pos).to(dev)
    edges = torch.from_numpy(small.edges).to(dev)
    e2c = torch.from_numpy(small.edge2cluster).to(dev)
    target = adjoint.density_target(pos, edges, e2c, small.num_clusters, (small.H, small.W))
    fit = adjoint.CardDensityFit.from_cards(small.cards, small.mvp, target, lr=1e-3)
    rep = fit.run(30)
    assert np.isfinite(rep.loss_last)
    assert rep.loss_last < 0.9 * rep.loss_first, (rep.loss_first, rep.loss_last)


def test_density_target_matches_oracle(dev, small):
    report("adjoint: strand-density target (popcount + 3x3 box) vs oracle mask")
    pos = torch.from_numpy(small.pos).to(dev)
    target = adjoint.density_target(pos, torch.from_numpy(small.edges).to(dev),
                                    torch.from_numpy(small.edge2cluster).to(dev),
                                    small.num_clusters, (small.H, small.W)).cpu().numpy()
    bits = oracle.dda_cluster_mask((small.H, small.W), small.pos, small.edges, small.edge2cluster,
                                   small.num_clusters, threads=8)
    pc = np.unpackbits(bits.view(np.uint8), axis=-1).sum(-1).astype(np.float64)
    pad = np.pad(pc, ((0, 0), (1, 1), (1, 1)))
    box = sum(pad[:, i:i + small.H, j:j + small.W] for i in range(3) for j in range(3)) / 9.0
    np.testing.assert_allclose(target, box, rtol=1e-6, atol=1e-6)


def test_errors(dev):
    pos = torch.zeros((1, 3, 4), device=dev)
    tri = torch.zeros((1, 3), dtype=torch.int32, device=dev)
    with pytest.raises(TypeError):
        adjoint.card_coverage(pos.double(), tri, (8, 8))
    with pytest.raises(ValueError):
        adjoint.card_coverage(pos.cpu(), tri.cpu(), (8, 8))
    with pytest.raises(ValueError):
        adjoint.card_coverage(pos[..., :3].contiguous(), tri, (8, 8))


@pytest.mark.slow
def test_full_adjoint_workload(dev):
    """1,024 cards x 20 quads over 8 views at 512^2: central differences on
    64 random vertices, then 200 Adam steps at lr 1e-3 on the strand
    density."""
    report("adjoint: building 50k-strand workload, 8 views 512^2")
    wl = synth.adjoint_workload()
    _check_fd(_fd_check(dev, wl, 8, 64, 7))
    pos = torch.from_numpy(wl.pos).to(dev)
    target = adjoint.density_target(pos, torch.from_numpy(wl.edges).to(dev),
                                    torch.from_numpy(wl.edge2cluster).to(dev), wl.num_clusters,
                                    (wl.H, wl.W))
    fit = adjoint.CardDensityFit.from_cards(wl.cards, wl.mvp, target, lr=1e-3)
    rep = fit.run(200)
    report(f"  200 steps: loss {rep.loss_first:.5g} -> {rep.loss_last:.5g}, "
           f"{rep.ms_per_step:.3f} ms/step")
    assert rep.loss_last < rep.loss_first
