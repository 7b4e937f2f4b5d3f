"""Silhouette-edge marking and the antialiased cluster-bitset loss.

Mirror of the reference module pystrns2cards/antialias.py.
"""
from typing import Any, Tuple

import torch

from . import _core
from .bitset_utils import popcount_bitset
from .dda import dda_compute_fragments


def mark_discontinuity_edges(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    tri: torch.Tensor,
    edges: torch.Tensor,
    edge2tri: torch.Tensor,
) -> Tuple[torch.Tensor, torch.Tensor]:
    """Silhouette / boundary edges per view (reference: antialias.py:8-52).

    Returns prim_ids (N,) int32 of ``b * E + e`` for edges whose two adjacent
    triangles lie on the same side in screen space (boundary edges always
    qualify), in ascending order, and normals (B*E, 2) float32: the unit
    screen normal of every edge, oriented away from its triangles.
    """
    H, W = resolution
    device = pos.device
    B = pos.shape[0]
    E = edges.shape[0]
    prim_ids = torch.full((B * E,), -1, dtype=torch.int32, device=device)
    normals = torch.zeros((B * E, 2), dtype=torch.float32, device=device)
    _core.mark_discontinuity_edges(H, W, tri.contiguous(), edges.contiguous(),
                                   edge2tri.contiguous(), pos.contiguous(), prim_ids, normals)
    out = torch.empty_like(prim_ids)
    n = _core.compact_valid_ints(prim_ids.contiguous(), out)
    return out[:n], normals


class _AntialiasedClusterBitsetLossFn(torch.autograd.Function):
    """Carries a precomputed d(loss)/d(pos) (reference: antialias.py:55-64)."""

    @staticmethod
    def forward(ctx: Any, loss_value: torch.Tensor, pos: torch.Tensor,
                d_pos: torch.Tensor) -> torch.Tensor:
        ctx.save_for_backward(d_pos)
        return loss_value

    @staticmethod
    def backward(ctx: Any, grad_output: torch.Tensor):
        d_pos, = ctx.saved_tensors
        return None, grad_output * d_pos, None


def antialiased_cluster_bitset_loss(
    bitset: torch.Tensor,
    target_bitset: torch.Tensor,
    pos: torch.Tensor,
    tri: torch.Tensor,
    edges: torch.Tensor,
    edge2tri: torch.Tensor,
    edge2cluster: torch.Tensor,
    kernel_radius: float = 7.0,
    rho: float = 1.0,
    count_xor_error: bool = False,
    *,
    fused: bool = False,
) -> torch.Tensor:
    """Antialiased silhouette loss between a rendered and a target cluster
    bitset (reference: antialias.py:66-130).

    The gradient with respect to ``pos`` is the edge-sampling estimate of
    antialias.cu: silhouette edges are DDA-rasterized, and at each pixel-
    centre crossing the Gaussian-weighted (sigma = R/4) residual of the
    edge's cluster over a (2*floor(R)+1)^2 window, false positives weighted
    by ``rho``, is pushed along the edge normal.  The value is 1, or the XOR
    popcount of the two bitsets when ``count_xor_error`` is set.

    ``fused=True`` (no reference counterpart) computes the same gradient
    without compacting the marked edges or materializing the fragment list:
    the backward kernel DDA-walks the marked edges itself, so the call makes
    no host synchronization (d_pos agrees up to the order of float atomics).
    """
    device = bitset.device
    assert target_bitset.device == device
    assert pos.device == device
    assert tri.device == device
    assert edges.device == device
    assert edge2tri.device == device
    assert edge2cluster.device == device

    _, H, W, _ = target_bitset.shape
    d_pos = torch.zeros_like(pos)
    if fused:
        B, E = pos.shape[0], edges.shape[0]
        marked = torch.full((B * E,), -1, dtype=torch.int32, device=device)
        normals = torch.zeros((B * E, 2), dtype=torch.float32, device=device)
        _core.mark_discontinuity_edges(H, W, tri.contiguous(), edges.contiguous(),
                                       edge2tri.contiguous(), pos.contiguous(), marked, normals)
        _core.backward_antialiased_cluster_bitset_edges(
            marked, normals, bitset.contiguous(), target_bitset.contiguous(), pos.contiguous(),
            edges.contiguous(), edge2cluster.contiguous(), d_pos, float(kernel_radius),
            float(rho))
    else:
        prim_ids, normals = mark_discontinuity_edges((H, W), pos, tri, edges, edge2tri)
        frag_pix, frag_attrs = dda_compute_fragments((H, W), pos, edges, prim_ids)
        _core.backward_antialiased_cluster_bitset(
            frag_pix.contiguous(), frag_attrs.contiguous(), bitset.contiguous(),
            target_bitset.contiguous(), pos.contiguous(), edges.contiguous(),
            normals.contiguous(), edge2cluster.contiguous(), d_pos, float(kernel_radius),
            float(rho))

    if count_xor_error:
        loss_value = popcount_bitset(torch.bitwise_xor(bitset, target_bitset)).sum()
    else:
        loss_value = torch.ones([], dtype=pos.dtype, device=device)
    return _AntialiasedClusterBitsetLossFn.apply(loss_value, pos, d_pos)
