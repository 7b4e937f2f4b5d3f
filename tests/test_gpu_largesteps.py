"""GPU parity of the grid-mesh large-steps system (to_differential /
from_differential on GridSystem) against the float64 oracle.

Tolerances, stated per test:
* apply (u = M x): elementwise |u - u_ref| <= 1e-6 * (|M| |x|), the
  rounding bound of a handful of fp32 operations per entry;
* solve (x = M^{-1} u): normwise backward error ||M x - u|| / (||M|| ||x||)
  <= 1e-6 (||M||_inf standing for ||M||_2), and forward error ||x - x_ref|| / ||x_ref|| <= 1e-6 + 3e-8 *
  kappa with kappa = 1 + 4 lam (1 + w_weight) >= cond(M) (the reference
  solves in fp32 too; its Cholesky has the same kind of bound).
"""
import numpy as np
import pytest
import torch

import oracle
from strns2cards_amd import optimizer
from strns2cards_amd.optimizer import GridSystem, from_differential, to_differential

pytestmark = pytest.mark.gpu

RAGGED = [(1, 1), (1, 2), (2, 1), (2, 2), (3, 7), (5, 40), (2, 300), (4, 17), (7, 3), (2, 64)]
CASES = [
    ("ragged", RAGGED, 100.0, 0.01),
    ("strips", [(2, 64)] * 37, 100.0, 0.01),
    ("identity", [(3, 9), (2, 5)], 0.0, 0.01),
    ("isotropic", [(4, 30), (2, 50)], 1.0, 1.0),
    ("stiff", [(2, 200), (6, 20)], 1000.0, 0.01),
    ("wide", [(21, 3), (30, 2)], 100.0, 0.5),
    ("long", [(2, 1365), (1, 1)], 100.0, 0.01),
]


def _system(shapes, lam, ww, dev):
    V, h, w = oracle.combined_grid_edges(shapes)
    M = oracle.largesteps_matrix(V, h, w, lam, ww)
    sys_ = GridSystem.from_edges(V, h, w, lam, ww, device=dev)
    assert list(zip(sys_.rows_host.tolist(), sys_.cols_host.tolist())) == list(shapes)
    return sys_, M


@pytest.mark.parametrize("name,shapes,lam,ww", CASES, ids=[c[0] for c in CASES])
def test_apply_matches_oracle(dev, name, shapes, lam, ww):
    print(f"[grid apply] {name}", flush=True)
    sys_, M = _system(shapes, lam, ww, dev)
    x = np.random.default_rng(3).standard_normal((sys_.num_verts, 3)).astype(np.float32)
    got = to_differential(sys_, torch.from_numpy(x).to(dev)).cpu().numpy()
    want = oracle.to_differential(M, x)
    bound = 1e-6 * (np.abs(M) @ np.abs(x.astype(np.float64)))
    assert np.all(np.abs(got - want) <= bound + 1e-30)


@pytest.mark.parametrize("name,shapes,lam,ww", CASES, ids=[c[0] for c in CASES])
def test_solve_matches_oracle(dev, name, shapes, lam, ww):
    print(f"[grid solve] {name}", flush=True)
    sys_, M = _system(shapes, lam, ww, dev)
    rng = np.random.default_rng(4)
    kappa = 1 + 4 * lam * (1 + ww)
    m_norm = np.abs(M).sum(1).max()  # ||M||_inf >= ||M||_2 (M symmetric)
    rough = rng.standard_normal((sys_.num_verts, 3))
    smooth = np.cumsum(rng.standard_normal((sys_.num_verts, 3)), 0) * 0.01
    for u in (rough.astype(np.float32), (M @ smooth).astype(np.float32)):
        x = from_differential(sys_, torch.from_numpy(u).to(dev)).cpu().numpy().astype(np.float64)
        want = oracle.from_differential(M, u)
        bwd = np.linalg.norm(M @ x - u) / (m_norm * np.linalg.norm(x))
        fwd = np.linalg.norm(x - want) / np.linalg.norm(want)
        assert bwd <= 1e-6, (name, bwd)
        assert fwd <= 1e-6 + 3e-8 * kappa, (name, fwd)


def test_solve_gradient(dev):
    """d/du of <g, M^{-1} u> is M^{-1} g (M symmetric), and the backward of
    to_differential is M g."""
    shapes = [(2, 33), (3, 5), (1, 4)]
    sys_, M = _system(shapes, 100.0, 0.01, dev)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((sys_.num_verts, 3)).astype(np.float32)
    u = torch.from_numpy(rng.standard_normal((sys_.num_verts, 3)).astype(np.float32)).to(dev)
    u.requires_grad_()
    (from_differential(sys_, u) * torch.from_numpy(g).to(dev)).sum().backward()
    want = oracle.from_differential(M, g)
    got = u.grad.cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-6 + 3e-8 * 401
    x = torch.zeros_like(u).requires_grad_()
    (to_differential(sys_, x) * torch.from_numpy(g).to(dev)).sum().backward()
    bound = 1e-6 * (np.abs(M) @ np.abs(g.astype(np.float64)))
    assert np.all(np.abs(x.grad.cpu().numpy() - M @ g) <= bound + 1e-30)


def test_round_trip_full_size(dev):
    """Card-fitting size: 16384 strips of 2 x 128 vertices (4.2M vertices).
    from_differential(to_differential(x)) returns x (size-independent check;
    each grid's error is bounded by kappa times fp32 rounding)."""
    print("[grid round trip] 16384 x (2, 128)", flush=True)
    sys_ = GridSystem([2] * 16384, [128] * 16384, 100.0, 0.01, device=dev)
    g = torch.Generator(device=dev).manual_seed(6)
    x = torch.randn(sys_.num_verts, 3, device=dev, generator=g)
    x2 = from_differential(sys_, to_differential(sys_, x))
    err = ((x2 - x).norm() / x.norm()).item()
    assert err <= 1e-6 + 3e-8 * 405, err
    worst = ((x2 - x).norm(dim=1).max() / x.norm(dim=1).max()).item()
    assert worst <= 1e-5, worst


def test_layout_limits(dev):
    sys_ = GridSystem([2], [1366], device=dev)  # 24 * 2732 + 16 > 65536 bytes of LDS
    with pytest.raises(RuntimeError, match="LDS"):
        sys_.solve(torch.zeros(sys_.num_verts, 3, device=dev))
    with pytest.raises(ValueError):
        sys_.solve(torch.zeros(sys_.num_verts + 1, 3, device=dev))
    empty = GridSystem([], [], device=dev)
    assert from_differential(empty, torch.zeros(0, 3, device=dev)).shape == (0, 3)
    assert optimizer.to_differential(empty, torch.zeros(0, 3, device=dev)).shape == (0, 3)


@pytest.mark.parametrize("name,shapes,lam,ww", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_grid_laplacian_matches_oracle(dev, name, shapes, lam, ww):
    """(L_h + w L_w) x against the oracle's laplacian_edges, elementwise
    within 1e-6 * (|L| |x|); the backward is L applied to the gradient."""
    V, h, w = oracle.combined_grid_edges(shapes)
    L = oracle.laplacian_edges(V, h) + ww * oracle.laplacian_edges(V, w)
    sys_ = GridSystem.from_edges(V, h, w, lam, ww, device=dev)
    x = np.random.default_rng(8).standard_normal((V, 3)).astype(np.float32)
    xt = torch.from_numpy(x).to(dev).requires_grad_(True)
    y = optimizer.grid_laplacian(sys_, xt, ww)
    bound = 1e-6 * (np.abs(L) @ np.abs(x.astype(np.float64)))
    assert np.all(np.abs(y.detach().cpu().numpy() - L @ x) <= bound + 1e-30)
    g = np.random.default_rng(9).standard_normal((V, 3)).astype(np.float32)
    (y * torch.from_numpy(g).to(dev)).sum().backward()
    bound = 1e-6 * (np.abs(L) @ np.abs(g.astype(np.float64)))
    assert np.all(np.abs(xt.grad.cpu().numpy() - L @ g) <= bound + 1e-30)



def test_solve_reference_default_size(dev):
    """2_texture.py's default strips (600 crossed 2-row strips sharing
    15,000 quads, 31,200 vertices) against a float64 sparse LU of the same
    M: forward error within 1e-6 + 3e-8 kappa."""
    rng = np.random.default_rng(11)
    length = rng.lognormal(0.0, 0.5, 300)
    quads = np.maximum(1, np.rint(7500 * length / length.sum())).astype(int)
    shapes = [(2, int(q) + 1) for q in np.repeat(quads, 2)]
    V, h, w = oracle.combined_grid_edges(shapes)
    sys_ = GridSystem([s[0] for s in shapes], [s[1] for s in shapes], 100.0, 0.01, device=dev)
    u = rng.standard_normal((V, 3)).astype(np.float32)
    x = from_differential(sys_, torch.from_numpy(u).to(dev)).cpu().numpy().astype(np.float64)
    want = oracle.from_differential_sparse(V, h, w, 100.0, 0.01, u)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) <= 1e-6 + 3e-8 * 405


def test_grid_quad_diagonals(dev):
    """(p0 + p2) - (p1 + p3) per quad, bit-exact against numpy on the
    oracle's grid-mesh quads (same float32 operations); the backward equals
    the adjoint scatter of the quads (float64, within 1e-6 * sum |dD|)."""
    shapes = RAGGED + [(3, 2), (2, 3)]
    V = sum(S * N for S, N in shapes)
    quads, off = [], 0
    for S, N in shapes:
        quads.append(oracle.grid_mesh(S, N)["quads"] + off)
        off += S * N
    quads = np.concatenate(quads)
    sys_ = GridSystem([s[0] for s in shapes], [s[1] for s in shapes], device=dev)
    pos = np.random.default_rng(12).standard_normal((V, 3)).astype(np.float32)
    pt = torch.from_numpy(pos).to(dev).requires_grad_(True)
    D = optimizer.grid_quad_diagonals(sys_, pt)
    want = (pos[quads[:, 0]] + pos[quads[:, 2]]) - (pos[quads[:, 1]] + pos[quads[:, 3]])
    np.testing.assert_array_equal(D.detach().cpu().numpy(), want)
    g = np.random.default_rng(13).standard_normal(want.shape).astype(np.float32)
    (D * torch.from_numpy(g).to(dev)).sum().backward()
    adj = np.zeros((V, 3))
    for k, sgn in enumerate((1.0, -1.0, 1.0, -1.0)):
        np.add.at(adj, quads[:, k], sgn * g.astype(np.float64))
    bound = np.zeros((V, 3))
    for k in range(4):
        np.add.at(bound, quads[:, k], np.abs(g.astype(np.float64)))
    assert np.all(np.abs(pt.grad.cpu().numpy() - adj) <= 1e-6 * bound + 1e-30)


@pytest.mark.parametrize("shape", [(31200, 3), (7, 3), (5000, 2), (40,)])
def test_uniform_adam_fused_matches_torch_ops(dev, shape):
    """UniformAdam(fused=True) (two HIP launches) against the torch-op step
    that restates optimizer.py:33-68 op for op, over 6 steps with a changing
    learning rate: parameters and both moments within 2e-6 normwise (the
    kernel rounds each product separately, torch may contract them); a 1-D
    parameter is one row, as the reference's p[..., :1] state makes it."""
    print(f"[uniform adam] shape={shape}", flush=True)
    from strns2cards_amd.optimizer import UniformAdam
    g = torch.Generator(device="cpu").manual_seed(sum(shape))
    x0 = torch.randn(shape, generator=g)
    pa = x0.clone().to(dev).requires_grad_(True)
    pb = x0.clone().to(dev).requires_grad_(True)
    oa = UniformAdam([pa], lr=0.1, betas=(0.9, 0.999))
    ob = UniformAdam([pb], lr=0.1, betas=(0.9, 0.999), fused=True)
    for it in range(6):
        grad = (torch.randn(shape, generator=g) * (1 + it)).to(dev)
        pa.grad = grad.clone()
        pb.grad = grad.clone()
        oa.step()
        ob.step()
        for o in (oa, ob):
            o.param_groups[0]["lr"] *= 0.9
    for a, b in ((pa, pb), (oa.state[pa]["g1"], ob.state[pb]["g1"]),
                 (oa.state[pa]["g2"], ob.state[pb]["g2"])):
        assert a.shape == b.shape
        err = ((a - b).abs().max() / a.abs().max()).item()
        assert err <= 2e-6, err
    assert ob.state[pb]["step"] == 6


def test_uniform_adam_fused_nan_and_errors(dev):
    """A NaN gradient entry makes the shared denominator NaN (torch's max
    propagates it), so every parameter turns NaN in both forms; CPU
    parameters are refused by the fused form."""
    from strns2cards_amd.optimizer import UniformAdam
    for fused in (False, True):
        p = torch.ones((10, 3), device=dev, requires_grad=True)
        opt = UniformAdam([p], lr=0.1, fused=fused)
        grad = torch.ones((10, 3), device=dev)
        grad[4, 1] = float("nan")
        p.grad = grad
        opt.step()
        assert torch.isnan(p).all(), fused
    q = torch.ones((4, 3), requires_grad=True)
    q.grad = torch.ones((4, 3))
    with pytest.raises(ValueError):
        UniformAdam([q], fused=True).step()
