"""CPU pins for the frame / rotation helpers and the Bishop-frame oracle.

minimal_rotation and tangent_frame are torch-only (curve_utils.py:9-102) and
run on host tensors; they are pinned by what a rotation must satisfy (maps
the source direction onto the target, orthogonal, det +1) and by the
reference's special cases.  The oracle's Bishop recurrence is pinned by the
properties of rotation-minimizing frames: orthonormal right-handed frames,
constant frames on a straight line, and on a planar curve a constant
out-of-plane component (the plane normal is transported unchanged).
"""
import numpy as np
import torch

import oracle
from strns2cards_amd import curve_utils as cu


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_minimal_rotation_maps_source_to_target():
    g = torch.Generator().manual_seed(0)
    a = torch.randn(200, 3, generator=g)
    b = torch.randn(200, 3, generator=g)
    R = cu.minimal_rotation(a, b).double()
    an = a.double() / a.double().norm(dim=-1, keepdim=True)
    bn = b.double() / b.double().norm(dim=-1, keepdim=True)
    assert torch.allclose((R @ an[..., None])[..., 0], bn, atol=1e-5)
    eye = torch.eye(3, dtype=torch.float64).expand_as(R)
    assert torch.allclose(R @ R.transpose(-1, -2), eye, atol=1e-5)
    assert torch.allclose(torch.linalg.det(R), torch.ones(200, dtype=torch.float64), atol=1e-5)
    # the rotation axis is the common normal: it is left fixed
    axis = torch.cross(an, bn, dim=-1)
    assert torch.allclose((R @ axis[..., None])[..., 0], axis, atol=1e-5)


def test_minimal_rotation_special_cases_and_broadcast():
    t = torch.tensor([[0.3, -0.2, 0.9]])
    assert torch.equal(cu.minimal_rotation(t, 2 * t)[0], torch.eye(3))
    flip = cu.minimal_rotation(t, -t)[0]
    assert torch.equal(flip, torch.diag(torch.tensor([-1.0, -1.0, 1.0])))
    src = torch.randn(4, 7, 3)
    R = cu.minimal_rotation(src, torch.tensor([0.0, 1.0, 0.0]))
    assert R.shape == (4, 7, 3, 3)


def test_tangent_frame_is_right_handed_orthonormal():
    g = torch.Generator().manual_seed(1)
    t = torch.randn(500, 3, generator=g)
    t[0] = torch.tensor([0.0, 0.0, -1.0])  # the branch point z = -1
    t[1] = torch.tensor([0.0, 0.0, 1.0])
    t[2] = torch.tensor([1.0, 0.0, 0.0])   # z = 0 takes the + branch
    n, b = cu.tangent_frame(t)
    tu = t / t.norm(dim=-1, keepdim=True)
    for u, v in ((n, b), (n, tu), (b, tu)):
        assert torch.allclose((u * v).sum(-1), torch.zeros(500), atol=1e-6)
    for u in (n, b):
        assert torch.allclose(u.norm(dim=-1), torch.ones(500), atol=1e-6)
    assert torch.allclose(torch.cross(n, b, dim=-1), tu, atol=1e-6)
    assert torch.allclose(n[2], torch.tensor([0.0, 0.0, -1.0]))


def _helix(n=64, turns=1.5):
    s = np.linspace(0, 2 * np.pi * turns, n)
    return np.stack([np.cos(s), np.sin(s), 0.3 * s], 1)


def test_oracle_bishop_orthonormal_and_tangent():
    x = np.stack([_helix(), _helix()[:, [1, 2, 0]] * 0.5]).astype(np.float32)
    T, N, B = oracle.bishop_frame(x)
    for u, v in ((T, N), (T, B), (N, B)):
        assert np.abs((u * v).sum(-1)).max() < 1e-5
    for u in (T, N, B):
        assert np.abs(np.linalg.norm(u, axis=-1) - 1).max() < 1e-5
    assert np.abs(np.cross(T, N) - B).max() < 1e-5
    # tangent = normalized sum of the adjacent segments
    d = np.diff(x.astype(np.float64), axis=1)
    want = np.zeros_like(x, dtype=np.float64)
    want[:, :-1] += d
    want[:, 1:] += d
    assert np.abs(T - unit(want)).max() < 1e-6


def test_oracle_bishop_straight_line_is_constant():
    x = (np.arange(20)[:, None] * np.array([0.2, -0.5, 0.7]))[None].astype(np.float32)
    T, N, B = oracle.bishop_frame(x)
    assert np.abs(N - N[:, :1]).max() < 1e-6
    assert np.abs(B - B[:, :1]).max() < 1e-6
    n0, b0 = cu.tangent_frame(torch.from_numpy(T[:, 0]))
    assert np.abs(N[:, 0] - n0.numpy()).max() < 1e-6
    assert np.abs(B[:, 0] - b0.numpy()).max() < 1e-6


def test_oracle_bishop_planar_curve_keeps_out_of_plane_component():
    th = np.linspace(0, 3.0, 40)
    x = np.stack([np.cos(th) + 0.2 * th, np.sin(2 * th), 0 * th], 1)
    rot = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))[0]
    x = (x @ rot.T)[None].astype(np.float32)
    plane_n = rot[:, 2]
    T, N, B = oracle.bishop_frame(x)
    nz = N[0] @ plane_n
    bz = B[0] @ plane_n
    assert np.abs(nz - nz[0]).max() < 2e-5
    assert np.abs(bz - bz[0]).max() < 2e-5


def test_oracle_bishop_is_rotation_minimizing():
    """The frame does not twist about the tangent: consecutive normals differ
    by a rotation whose tangential component vanishes as the sampling
    refines (the double reflection is exact on circular arcs)."""
    twist = []
    for n in (50, 200, 800):
        x = _helix(n)[None].astype(np.float64).astype(np.float32)
        T, N, B = oracle.bishop_frame(x)
        # component of the normal's change along the previous binormal
        twist.append(np.abs((N[0, 1:] * B[0, :-1]).sum(-1)).max())
    assert twist[2] < twist[0]
    assert twist[2] < 2e-3
