"""Differentiable card coverage with a hand-written adjoint.

The reference fits cards by differentiating its rasterize + antialias chain
(3_fitting.py:380-400: ``render_bitset_from_triangles`` ->
``antialiased_cluster_bitset_loss``), whose gradient is a surrogate built
along silhouette edges (antialias.cu).  This module gives that chain a form
whose gradient is exact: the rendered image is the area-sampled coverage

    C[b, h, w] = sum_t area(triangle t of view b ∩ [w, w+1] x [h, h+1])

in the reference's screen coordinates (dda.cu:33-36), which is continuous
in the vertices.  ``CardCoverage`` runs it forward in one HIP kernel and
backward in another; the backward recomputes every triangle's screen
geometry and each pixel's clip from the saved clip-space vertices (nothing
per pixel is stored between the passes) and applies the transport theorem
along the triangle edges inside each pixel box.

``CardDensityFit`` is the fitting problem this form exists for: card
centre-lines and widths as fp32 parameters, rendered over several
orthographic views, L2 against the clusters' DDA strand density
(popcount of the DDA cluster mask, box-filtered 3x3), Adam.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

from . import _core
from .bitset_utils import popcount_bitset
from .dda import dda_cluster_mask

__all__ = ["CardCoverage", "card_coverage", "strand_density", "density_target", "CardDensityFit",
           "DensityFitReport"]


def _check(pos: torch.Tensor, tri: torch.Tensor) -> None:
    if not pos.is_cuda or not tri.is_cuda:
        raise ValueError("pos and tri must be on CUDA")
    if pos.dtype != torch.float32 or tri.dtype != torch.int32:
        raise TypeError("pos must be float32 and tri int32")
    if pos.dim() != 3 or pos.shape[-1] != 4:
        raise ValueError("pos must have shape (B, V, 4)")
    if tri.dim() != 2 or tri.shape[-1] != 3:
        raise ValueError("tri must have shape (T, 3)")


class CardCoverage(torch.autograd.Function):
    """(B, V, 4) clip-space vertices -> (B, H, W) area-sampled coverage."""

    @staticmethod
    def forward(ctx, pos: torch.Tensor, tri: torch.Tensor, H: int, W: int) -> torch.Tensor:
        pos = pos.contiguous()
        tri = tri.contiguous()
        cov = torch.zeros((pos.shape[0], H, W), dtype=torch.float32, device=pos.device)
        _core.card_coverage_forward(pos, tri, cov)
        ctx.save_for_backward(pos, tri)
        return cov

    @staticmethod
    def backward(ctx, d_cov: torch.Tensor):
        pos, tri = ctx.saved_tensors
        d_pos = torch.zeros_like(pos)
        _core.card_coverage_backward(pos, tri, d_cov.contiguous(), d_pos)
        return d_pos, None, None, None


def card_coverage(pos: torch.Tensor, tri: torch.Tensor, resolution) -> torch.Tensor:
    """Area-sampled coverage of triangles ``tri`` (T, 3) over clip-space
    vertices ``pos`` (B, V, 4); differentiable in ``pos``."""
    _check(pos, tri)
    H, W = resolution
    return CardCoverage.apply(pos, tri, int(H), int(W))


def strand_density(bitset: torch.Tensor) -> torch.Tensor:
    """(B, H, W) number of clusters whose strands cover each pixel (popcount
    of the DDA cluster mask), box-filtered 3 x 3 with zero padding."""
    d = popcount_bitset(bitset)
    return F.avg_pool2d(d[:, None], 3, stride=1, padding=1, count_include_pad=True)[:, 0]


@dataclass
class DensityFitReport:
    steps: int
    loss_first: float
    loss_last: float
    ms_per_step: float
    losses: list


class CardDensityFit:
    """Cards (strips of Q quads) fitted to the strand density of their
    clusters over B orthographic views.

    Parameters are the strip centre-lines ``centre`` (K, Q+1, 3) and the
    per-row widths ``width`` (K, Q+1); the side direction of every row is
    fixed, so a row's two vertices are centre -/+ width/2 * side (the vertex
    order of the card strips: row 0 then row 1, each Q+1 long)."""

    def __init__(self, centre, width, side, tri, mvp, target, lr: float = 1e-3):
        dev = target.device
        self.centre = torch.nn.Parameter(torch.as_tensor(centre, dtype=torch.float32, device=dev))
        self.width = torch.nn.Parameter(torch.as_tensor(width, dtype=torch.float32, device=dev))
        self.side = torch.as_tensor(side, dtype=torch.float32, device=dev)
        self.tri = torch.as_tensor(tri, dtype=torch.int32, device=dev).contiguous()
        self.mvpT = torch.as_tensor(mvp, dtype=torch.float32, device=dev).transpose(1, 2).contiguous()
        self.target = target
        self.H, self.W = target.shape[1:]
        self.opt = torch.optim.Adam([self.centre, self.width], lr=lr)

    @classmethod
    def from_cards(cls, cards, mvp, target, lr: float = 1e-3):
        """Parameters from card strips built as row0 = c - w/2 s, row1 = c + w/2 s."""
        v = np.asarray(cards.verts, np.float64)
        K = (v.shape[0] - cards.tri.shape[0]) // 2  # V = 2(Q+1)K, T = 2QK
        R = v.shape[0] // (2 * K)
        rows = v.reshape(K, 2, R, 3)
        centre = rows.mean(1)
        off = rows[:, 1] - rows[:, 0]
        width = np.linalg.norm(off, axis=-1)
        side = off / np.maximum(width[..., None], 1e-12)
        return cls(centre, width, side, cards.tri, mvp, target, lr)

    def verts(self) -> torch.Tensor:
        half = 0.5 * self.width[..., None] * self.side
        rows = torch.stack([self.centre - half, self.centre + half], 1)  # (K, 2, R, 3)
        return rows.reshape(-1, 3)

    def clip(self, verts: torch.Tensor) -> torch.Tensor:
        vh = torch.cat([verts, torch.ones_like(verts[:, :1])], -1)
        return torch.matmul(vh[None], self.mvpT)  # (B, V, 4)

    def loss(self, verts: torch.Tensor | None = None) -> torch.Tensor:
        v = self.verts() if verts is None else verts
        cov = card_coverage(self.clip(v), self.tri, (self.H, self.W))
        return (cov - self.target).square().mean()

    def step(self) -> torch.Tensor:
        self.opt.zero_grad(set_to_none=True)
        loss = self.loss()
        loss.backward()
        self.opt.step()
        return loss.detach()

    def run(self, steps: int) -> DensityFitReport:
        losses = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            losses.append(self.step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ls = torch.stack(losses).cpu().tolist() if losses else []
        with torch.no_grad():
            last = float(self.loss())
        return DensityFitReport(steps, ls[0] if ls else last, last,
                                1e3 * dt / max(steps, 1), ls)


def density_target(pos: torch.Tensor, edges: torch.Tensor, edge2cluster: torch.Tensor,
                   num_clusters: int, resolution) -> torch.Tensor:
    """Strand-density target of ``CardDensityFit``: the DDA cluster mask of
    all strands in every view, popcounted and box-filtered 3 x 3."""
    bits = dda_cluster_mask(tuple(resolution), pos, edges, edge2cluster, num_clusters)
    return strand_density(bits)
