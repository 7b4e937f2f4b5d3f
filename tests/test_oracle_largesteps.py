"""Pins for the large-steps oracle and the grid-layout host logic (CPU).

The reference's solve lives in largesteps 0.2.2 / cholespy 2.1.0, which
are not installed, and it ships no vectors for them.  The oracle's
Laplacian, system matrix and solve are pinned by hand-computed matrices,
the closed-form spectrum of a grid's system matrix (a swapped h/w weight,
a dropped term or a wrong sign changes it) and an independent dense solve.
"""
import numpy as np
import pytest
import torch

import oracle
from strns2cards_amd.optimizer import GridSystem, grid_edges


def test_wisp_edges_by_hand():
    h, w = oracle.wisp_grid_edges(2, 3)
    assert h.tolist() == [[0, 1], [1, 2], [3, 4], [4, 5]]
    assert w.tolist() == [[0, 3], [1, 4], [2, 5]]
    V, h2, w2 = oracle.combined_grid_edges([(1, 2), (2, 2)])
    assert V == 6
    assert h2.tolist() == [[0, 1], [2, 3], [4, 5]]
    assert w2.tolist() == [[2, 4], [3, 5]]


def test_laplacian_by_hand():
    L = oracle.laplacian_edges(3, [[0, 1], [1, 2]])
    assert L.tolist() == [[1, -1, 0], [-1, 2, -1], [0, -1, 1]]
    # repeated and reversed edges collapse (unique over the symmetrized pairs)
    L = oracle.laplacian_edges(2, [[0, 1], [1, 0], [0, 1]])
    assert L.tolist() == [[1, -1], [-1, 1]]
    # a self-loop's -1 and +1 land on the same diagonal entry
    L = oracle.laplacian_edges(3, [[0, 1], [2, 2]])
    assert L.tolist() == [[1, -1, 0], [-1, 1, 0], [0, 0, 0]]


@pytest.mark.parametrize("S,N,lam,ww", [(3, 5, 100.0, 0.01), (2, 7, 1.0, 1.0), (4, 4, 10.0, 0.3)])
def test_system_matrix_spectrum(S, N, lam, ww):
    V, h, w = oracle.combined_grid_edges([(S, N)])
    M = oracle.largesteps_matrix(V, h, w, lam, ww)
    nu = 2 - 2 * np.cos(np.pi * np.arange(N) / N)  # path of N vertices (h direction)
    mu = 2 - 2 * np.cos(np.pi * np.arange(S) / S)  # path of S vertices (w direction)
    want = np.sort((1 + lam * (nu[None, :] + ww * mu[:, None])).ravel())
    np.testing.assert_allclose(np.linalg.eigvalsh(M), want, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(M @ np.ones(V), np.ones(V), atol=1e-12)
    assert np.array_equal(M, M.T)


def test_solve_by_hand_and_dense():
    M = oracle.largesteps_matrix(2, [[0, 1]], np.zeros((0, 2)), 1.0, 0.5)
    assert M.tolist() == [[2, -1], [-1, 2]]
    x = oracle.from_differential(M, np.array([[1.0, 0, 0], [0, 3.0, 0]]))
    np.testing.assert_allclose(x, [[2 / 3, 1.0, 0], [1 / 3, 2.0, 0]], atol=1e-15)
    rng = np.random.default_rng(1)
    V, h, w = oracle.combined_grid_edges([(2, 9), (3, 4), (1, 5)])
    M = oracle.largesteps_matrix(V, h, w, 100.0, 0.01)
    u = rng.standard_normal((V, 3))
    x = oracle.from_differential(M, u)
    np.testing.assert_allclose(x, np.linalg.solve(M, u), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(oracle.to_differential(M, x), u, atol=1e-10)


# ------------------------------------------------------------ host logic


@pytest.mark.parametrize("shapes", [
    [(2, 5)], [(1, 1)], [(1, 1), (1, 1), (3, 1)], [(2, 4), (2, 4), (1, 4), (5, 2)],
    [(2, 30), (3, 7), (1, 9), (2, 1), (4, 4), (7, 2)],
])
def test_grid_layout_from_edges(shapes):
    V, h, w = oracle.combined_grid_edges(shapes)
    rng = np.random.default_rng(len(shapes))
    # any edge order and orientation
    hp = h[rng.permutation(len(h))][:, ::-1] if len(h) else h
    sys_ = GridSystem.from_edges(V, torch.from_numpy(hp.copy()), w[rng.permutation(len(w))],
                                 device="cpu")
    assert list(zip(sys_.rows_host.tolist(), sys_.cols_host.tolist())) == list(shapes)
    assert sys_.num_verts == V
    eh, ew = grid_edges(sys_.rows_host, sys_.cols_host)
    assert np.array_equal(eh, h) and np.array_equal(ew, w)


def test_grid_layout_rejects_other_meshes():
    V, h, w = oracle.combined_grid_edges([(3, 4)])
    with pytest.raises(ValueError):  # a missing w_edge
        GridSystem.from_edges(V, h, w[:-1], device="cpu")
    with pytest.raises(ValueError):  # a diagonal edge
        GridSystem.from_edges(V, h, np.concatenate([w, [[0, 5]]]), device="cpu")
    with pytest.raises(ValueError):  # a w_edge in the h list
        GridSystem.from_edges(V, np.concatenate([h, [[0, 4]]]), w, device="cpu")
    with pytest.raises(ValueError):
        GridSystem.from_edges(V, h, np.array([[0, V]]), device="cpu")
    with pytest.raises(ValueError):
        GridSystem([2, 0], [3, 3], device="cpu")
    with pytest.raises(ValueError):
        GridSystem([2], [3], lam=-1.0, device="cpu")


def test_sparse_solve_agrees_with_cholesky():
    V, h, w = oracle.combined_grid_edges([(2, 40), (3, 9), (1, 7)])
    M = oracle.largesteps_matrix(V, h, w, 100.0, 0.01)
    u = np.random.default_rng(2).standard_normal((V, 3))
    np.testing.assert_allclose(oracle.from_differential_sparse(V, h, w, 100.0, 0.01, u),
                               oracle.from_differential(M, u), rtol=1e-9, atol=1e-11)


def test_fused_uniform_adam_refuses_host_tensors():
    """UniformAdam(fused=True) checks its parameters before any library
    call (host tensors raise, the parameter is untouched); the torch-op
    form's first step matches its closed form."""
    from strns2cards_amd.optimizer import UniformAdam
    p = torch.ones((4, 3), requires_grad=True)
    p.grad = torch.ones((4, 3))
    opt = UniformAdam([p], lr=0.1, fused=True)
    with pytest.raises(ValueError, match="GPU"):
        opt.step()
    assert torch.equal(p.detach(), torch.ones((4, 3)))
    q = torch.ones((4, 3), requires_grad=True)
    q.grad = torch.ones((4, 3))
    UniformAdam([q], lr=0.1).step()  # the torch-op form runs anywhere
    # first step, unit gradient: m1 = 1 and m2 = 3 (the isotropic moment sums
    # the three channels), so each coordinate moves by lr / sqrt(3); a
    # per-channel Adam would move it by lr
    want = 1.0 - 0.1 / (np.sqrt(3.0) + 1e-8)
    assert torch.allclose(q.detach(), torch.full((4, 3), want), atol=1e-6)
