"""DDA rasterization of line meshes (strands) into fragments.

Mirror of the reference module pystrns2cards/dda.py: same functions,
arguments, returned layouts and exceptions, running on the gfx950 kernels of
``libstrns2cards_hip.so`` through ``_core``.
"""
from typing import Optional, Tuple

import torch

from . import _core


def dda_compute_fragments(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    edges: torch.Tensor,
    edge_ids: Optional[torch.Tensor] = None,
) -> Tuple[torch.Tensor, torch.Tensor]:
    """Rasterize the edges of a line mesh with DDA (reference: dda.py:5-76).

    Args:
        resolution: (H, W).
        pos: (B, V, 4) float32 clip-space vertices on the GPU.
        edges: (E, 2) int32 vertex pairs on the GPU.
        edge_ids: optional (P,) int32 ids ``b * E + e`` selecting what to draw;
            defaults to every edge of every view.

    Returns:
        frag_pix (N, 3) int32 ``[b, h, w]`` and frag_attrs (N, 4) float32
        ``[axis, t1, z_ndc, e + 1]``, in primitive order.
    """
    H, W = resolution
    if not pos.is_cuda or not edges.is_cuda:
        raise ValueError("Input tensors must be on CUDA.")
    if pos.dtype != torch.float32 or edges.dtype != torch.int32:
        raise ValueError("`pos` must be float32 and `edges` must be int32.")
    if pos.ndim != 3 or pos.shape[2] != 4:
        raise ValueError("`pos` must have shape (B, V, 4).")
    if edges.ndim != 2 or edges.shape[1] != 2:
        raise ValueError("`edges` must have shape (E, 2).")
    if not pos.is_contiguous() or not edges.is_contiguous():
        raise ValueError("Input tensors must be contiguous.")

    B = pos.shape[0]
    E = edges.shape[0]
    device = pos.device
    if edge_ids is not None and (not edge_ids.is_cuda or edge_ids.dtype != torch.int32
                                 or edge_ids.ndim != 1):
        raise ValueError("`edge_ids` must be a 1D int32 CUDA tensor.")

    # The reference's span, -1-filled candidate list and compaction
    # (dda.py:45-76) as two passes over the surviving fragments: the same
    # fragments in the same order, with no candidate list (edge_ids None is
    # the reference's arange over every edge of every view).
    P = B * E if edge_ids is None else edge_ids.shape[0]
    valid_prefix = torch.empty((P,), dtype=torch.int32, device=device)
    n = _core.dda_count_fragments(edge_ids, H, W, pos, edges, valid_prefix)
    frag_pix = torch.empty((n, 3), dtype=torch.int32, device=device)
    frag_attrs = torch.empty((n, 4), dtype=torch.float32, device=device)
    _core.dda_emit_fragments(valid_prefix, edge_ids, H, W, pos, edges, frag_pix, frag_attrs)
    return frag_pix, frag_attrs


class DDAInterpolateAttributes(torch.autograd.Function):
    """Per-vertex attributes interpolated along DDA fragments, with an
    atomic scatter backward (reference: dda.py:79-120)."""

    @staticmethod
    def forward(ctx, rast: torch.Tensor, attr: torch.Tensor, edges: torch.Tensor) -> torch.Tensor:
        if not rast.is_cuda or not attr.is_cuda or not edges.is_cuda:
            raise ValueError("All input tensors must be CUDA tensors.")
        if rast.dtype != torch.float32 or attr.dtype != torch.float32:
            raise ValueError("`rast` and `attr` must be float32.")
        if edges.dtype != torch.int32:
            raise ValueError("`edges` must be int32.")
        if rast.ndim != 4 or rast.shape[-1] != 4:
            raise ValueError("`rast` must have shape (B, H, W, 4).")
        if attr.ndim != 2:
            raise ValueError("`attr` must have shape (V, C).")
        if edges.ndim != 2 or edges.shape[1] != 2:
            raise ValueError("`edges` must have shape (E, 2).")
        B, H, W, _ = rast.shape
        C = attr.shape[1]
        image = torch.zeros((B, H, W, C), dtype=torch.float32, device=rast.device)
        _core.dda_interpolate_attributes(rast.contiguous(), attr.contiguous(),
                                         edges.contiguous(), image)
        ctx.save_for_backward(rast, attr, edges)
        return image

    @staticmethod
    def backward(ctx, d_image: torch.Tensor):
        rast, attr, edges = ctx.saved_tensors
        d_attr = torch.zeros_like(attr)
        _core.backward_dda_interpolate_attributes(rast.contiguous(), d_image.contiguous(),
                                                  edges.contiguous(), d_attr)
        return None, d_attr, None


def dda_interpolate_attributes(rast: torch.Tensor, attr: torch.Tensor,
                               edges: torch.Tensor) -> torch.Tensor:
    """(B, H, W, C) image of attributes interpolated along DDA fragments
    (reference: dda.py:124-136)."""
    return DDAInterpolateAttributes.apply(rast, attr, edges)


def dda_cluster_mask(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    edges: torch.Tensor,
    edge2cluster: torch.Tensor,
    cluster_count: Optional[int] = None,
    tile_log2: Optional[int] = None,
) -> torch.Tensor:
    """Fused form of ``dda_compute_fragments`` followed by
    ``cluster_mask_from_fragments`` (the target-mask sequence of
    3_fitting.py:254-261): one device pass, no fragment list, no host sync.

    Returns a (B, H, W, ceil(K/32)) int32 bitset with B = pos.shape[0]; it
    equals the two-call result on every view (views that receive no fragment
    are all zero, where the two-call form drops trailing empty views).
    ``tile_log2`` fixes the binning tile side (2..6; 0 = automatic); by
    default the side ``autotune.tune_dda`` chose for this shape is used.
    The result does not depend on it.
    """
    H, W = resolution
    if not (pos.is_cuda and edges.is_cuda and edge2cluster.is_cuda):
        raise ValueError("Input tensors must be on CUDA.")
    if pos.dtype != torch.float32 or edges.dtype != torch.int32 or edge2cluster.dtype != torch.int32:
        raise ValueError("`pos` must be float32; `edges` and `edge2cluster` int32.")
    if pos.ndim != 3 or pos.shape[2] != 4:
        raise ValueError("`pos` must have shape (B, V, 4).")
    if edges.ndim != 2 or edges.shape[1] != 2:
        raise ValueError("`edges` must have shape (E, 2).")
    if edge2cluster.ndim != 1 or edge2cluster.shape[0] != edges.shape[0]:
        raise ValueError("`edge2cluster` must have shape (E,).")
    if cluster_count is None:
        cluster_count = int(edge2cluster.max().item()) + 1 if edge2cluster.numel() else 1
    S = (cluster_count + 31) // 32
    bitset = torch.empty((pos.shape[0], H, W, S), dtype=torch.int32, device=pos.device)
    if tile_log2 is None:
        from .autotune import dda_tile_for
        tile_log2 = dda_tile_for(pos.shape[0], H, W, S, edges.shape[0])
    _core.dda_cluster_mask(pos.contiguous(), edges.contiguous(), edge2cluster.contiguous(), bitset,
                           int(tile_log2))
    return bitset
