"""Triangle rasterization, depth test, attribute interpolation and cluster
bitset masks.

Mirror of the reference module pystrns2cards/rasterize.py on the gfx950
kernels of ``libstrns2cards_hip.so``.
"""
from typing import Optional, Tuple

import torch

from . import _core


def compute_fragments(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    tri: torch.Tensor,
) -> Tuple[torch.Tensor, torch.Tensor]:
    """All covered (view, pixel, triangle) samples (reference:
    rasterize.py:6-53).

    Returns frag_pix (N, 3) int32 ``[b, y, x]`` and frag_attrs (N, 4) float32
    ``[b0, b1, z_ndc, t + 1]`` in (b, t)-major, row-major pixel order.
    """
    H, W = resolution
    if not pos.is_cuda or not tri.is_cuda:
        raise ValueError("Input tensors must be on CUDA.")
    if pos.dtype != torch.float32 or tri.dtype != torch.int32:
        raise ValueError("`pos` must be float32 and `tri` must be int32.")
    if pos.ndim != 3 or pos.shape[2] != 4:
        raise ValueError("`pos` must have shape (B, V, 4).")
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("`tri` must have shape (T, 3).")
    if not pos.is_contiguous() or not tri.is_contiguous():
        raise ValueError("Input tensors must be contiguous.")

    B = pos.shape[0]
    T = tri.shape[0]
    device = pos.device
    rects = torch.empty((B * T, 4), dtype=torch.int32, device=device)
    prefix = torch.empty((B * T,), dtype=torch.int32, device=device)
    num_frags = _core.compute_triangle_rects(H, W, pos, tri, rects, prefix)

    frag_pix = torch.full((num_frags, 3), -1, dtype=torch.int32, device=device)
    frag_attrs = torch.empty((num_frags, 4), dtype=torch.float32, device=device)
    _core.compute_fragments(H, W, pos, tri, prefix, rects, frag_pix, frag_attrs)

    pix_out = torch.empty_like(frag_pix)
    attrs_out = torch.empty_like(frag_attrs)
    n = _core.filter_valid_fragments(frag_pix, frag_attrs, pix_out, attrs_out)
    return pix_out[:n], attrs_out[:n]


def depth_test(
    B: int,
    resolution: Tuple[int, int],
    frag_pix: torch.Tensor,
    frag_attrs: torch.Tensor,
) -> torch.Tensor:
    """Nearest fragment per pixel (reference: rasterize.py:55-95).

    The winner has the smallest key ``(bits(z + 2) << 32) | index``: smallest
    depth, ties to the lowest fragment index.  Returns (B, H, W, 4) float32
    with the winner's attributes, zeros where nothing is drawn.
    """
    if not (frag_pix.is_cuda and frag_attrs.is_cuda):
        raise ValueError("Input tensors must be on CUDA.")
    if frag_pix.dtype != torch.int32:
        raise ValueError("frag_pix must be int32.")
    if frag_attrs.dtype != torch.float32:
        raise ValueError("frag_attrs must be float32.")
    if frag_pix.ndim != 2 or frag_pix.shape[1] != 3:
        raise ValueError("frag_pix must have shape (N, 3).")
    if frag_attrs.ndim != 2 or frag_attrs.shape[1] != 4:
        raise ValueError("frag_attrs must have shape (N, 4).")
    if frag_pix.shape[0] != frag_attrs.shape[0]:
        raise ValueError("frag_pix and frag_attrs must match in length.")
    H, W = resolution
    rast = torch.zeros((B, H, W, 4), dtype=torch.float32, device=frag_pix.device)
    _core.depth_test(H, W, frag_pix.contiguous(), frag_attrs.contiguous(), rast)
    return rast


def rasterize(resolution: Tuple[int, int], pos: torch.Tensor, tri: torch.Tensor) -> torch.Tensor:
    """``compute_fragments`` then ``depth_test`` (reference: rasterize.py:97-116)."""
    frag_pix, frag_attrs = compute_fragments(resolution, pos, tri)
    return depth_test(pos.shape[0], resolution, frag_pix, frag_attrs)


class InterpolateTriangleAttributes(torch.autograd.Function):
    """Barycentric attribute interpolation with an atomic scatter backward
    (reference: rasterize.py:119-148)."""

    @staticmethod
    def forward(ctx, rast: torch.Tensor, attr: torch.Tensor, tri: torch.Tensor) -> torch.Tensor:
        if not (rast.is_cuda and attr.is_cuda and tri.is_cuda):
            raise ValueError("All inputs must be CUDA tensors.")
        B, H, W, _ = rast.shape
        C = attr.shape[1]
        image = torch.zeros((B, H, W, C), dtype=attr.dtype, device=attr.device)
        _core.interpolate_triangle_attributes(rast.contiguous(), attr.contiguous(),
                                              tri.contiguous(), image)
        ctx.save_for_backward(rast, tri, attr)
        return image

    @staticmethod
    def backward(ctx, grad_output: torch.Tensor):
        rast, tri, attr = ctx.saved_tensors
        d_attr = torch.zeros_like(attr)
        _core.backward_interpolate_triangle_attributes(rast.contiguous(), grad_output.contiguous(),
                                                       tri.contiguous(), d_attr)
        return None, d_attr, None


def interpolate_triangle_attributes(rast: torch.Tensor, attr: torch.Tensor,
                                    tri: torch.Tensor) -> torch.Tensor:
    """(B, H, W, C) barycentric interpolation (reference: rasterize.py:151-167)."""
    return InterpolateTriangleAttributes.apply(rast, attr, tri)


def cluster_mask_from_fragments(
    resolution: Tuple[int, int],
    frag_pix: torch.Tensor,
    frag_attrs: torch.Tensor,
    tri2cluster: torch.Tensor,
    cluster_count: Optional[int] = None,
) -> torch.Tensor:
    """Per-pixel bitset of the clusters whose primitives cover it
    (reference: rasterize.py:170-214).

    ``B`` is ``frag_pix[:, 0].max() + 1`` as in the reference (so an empty
    fragment list raises, and trailing empty views are dropped); cluster ``k``
    is bit ``k % 32`` of word ``k // 32``.
    """
    H, W = resolution
    if frag_pix.dtype != torch.int32 or frag_pix.ndim != 2 or frag_pix.shape[1] != 3:
        raise ValueError("frag_pix must be a (N, 3) int32 tensor.")
    if frag_attrs.dtype != torch.float32 or frag_attrs.ndim != 2 or frag_attrs.shape[1] != 4:
        raise ValueError("frag_attrs must be a (N, 4) float32 tensor.")
    if tri2cluster.dtype != torch.int32 or tri2cluster.ndim != 1:
        raise ValueError("tri2cluster must be a (T,) int32 tensor.")
    if not (frag_pix.is_cuda and frag_attrs.is_cuda and tri2cluster.is_cuda):
        raise ValueError("All input tensors must be CUDA tensors.")

    # the reference's frag_pix[:, 0].max() + 1, read with 16-byte loads over
    # the rows (torch's strided column copy and reduction took 4x longer)
    if frag_pix.shape[0] == 0 or not frag_pix.is_contiguous() or frag_pix.data_ptr() % 16:
        B = int(frag_pix[:, 0].max().item()) + 1
    else:
        B = _core.max_frag_view(frag_pix) + 1
    if cluster_count is None:
        cluster_count = int(tri2cluster.max().item()) + 1
    S = (cluster_count + 31) // 32
    # The reference's zeroed bitset and one atomic OR per fragment, as one
    # binned pass that writes every word (the same bits for fragments the
    # reference's kernel can address: pixels inside the image, cluster ids
    # below 32 * S).
    bitset = torch.empty((B, H, W, S), dtype=torch.int32, device=frag_pix.device)
    _core.fragment_cluster_mask(frag_pix.contiguous(), frag_attrs.contiguous(),
                                tri2cluster.contiguous(), bitset)
    return bitset


def triangle_cluster_mask(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    tri: torch.Tensor,
    tri2cluster: torch.Tensor,
    cluster_count: Optional[int] = None,
) -> torch.Tensor:
    """Fused form of ``compute_fragments`` + ``cluster_mask_from_fragments``
    (render_bitset_from_triangles, 3_fitting.py:154-163): one device pass with
    no fragment list and no host sync.  Returns (B, H, W, S) with
    B = pos.shape[0]."""
    H, W = resolution
    if not (pos.is_cuda and tri.is_cuda and tri2cluster.is_cuda):
        raise ValueError("Input tensors must be on CUDA.")
    if pos.dtype != torch.float32 or tri.dtype != torch.int32 or tri2cluster.dtype != torch.int32:
        raise ValueError("`pos` must be float32; `tri` and `tri2cluster` int32.")
    if pos.ndim != 3 or pos.shape[2] != 4:
        raise ValueError("`pos` must have shape (B, V, 4).")
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("`tri` must have shape (T, 3).")
    if cluster_count is None:
        cluster_count = int(tri2cluster.max().item()) + 1 if tri2cluster.numel() else 1
    S = (cluster_count + 31) // 32
    bitset = torch.empty((pos.shape[0], H, W, S), dtype=torch.int32, device=pos.device)
    _core.triangle_cluster_mask(pos.contiguous(), tri.contiguous(), tri2cluster.contiguous(),
                                bitset)
    return bitset
