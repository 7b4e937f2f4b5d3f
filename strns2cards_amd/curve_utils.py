"""Strand utilities: arclength, Gaussian smoothing and rotation skinning.

Mirror of the kernel-backed part of the reference module
pystrns2cards/curve_utils.py (curve_utils.py:167-240).
"""
import torch

from . import _core


def compute_arclength(strands: torch.Tensor) -> torch.Tensor:
    """(N, V) cumulative polyline length of (N, V, 3) strands, starting at 0
    (reference: curve_utils.py:167-180)."""
    N = strands.shape[0]
    seg = torch.norm(strands[:, 1:] - strands[:, :-1], dim=2)
    zero = torch.zeros((N, 1), device=strands.device, dtype=seg.dtype)
    return torch.cat([zero, torch.cumsum(seg, dim=1)], dim=1)


def smooth_strands(strands: torch.Tensor, beta: float) -> torch.Tensor:
    """Gaussian smoothing in arclength (reference: curve_utils.py:182-206).

    Each vertex becomes the average of its strand's vertices weighted by
    exp(-(s_i - s_j)^2 / beta^2), vertices farther than 3*beta in arclength
    excluded.
    """
    if not strands.is_cuda:
        raise ValueError("smooth_strands: Input tensor must be on CUDA.")
    if strands.ndim != 3 or strands.shape[-1] != 3:
        raise ValueError("smooth_strands: Input tensor must have shape (N, V, 3).")
    arclength = compute_arclength(strands)
    out = torch.zeros_like(strands)
    _core.smooth_strands(strands, arclength, out, beta)
    return out


def skinning(
    strands: torch.Tensor,
    handle_strands: torch.Tensor,
    canonical_handles: torch.Tensor,
    R: torch.Tensor,
    beta: float,
) -> torch.Tensor:
    """Gaussian-weighted rotation skinning (reference: curve_utils.py:209-240).

    x_i -> sum_j w_ij (R_j (x_i - h_j) + c_j) / sum_j w_ij with the
    smoothing weights above; a vertex with total weight <= 1e-8 keeps x_i.
    """
    if not (strands.is_cuda and handle_strands.is_cuda and canonical_handles.is_cuda and R.is_cuda):
        raise ValueError("skinning: All input tensors must be on CUDA.")
    N, V, _ = strands.shape
    if (handle_strands.shape != (N, V, 3) or canonical_handles.shape != (N, V, 3)
            or R.shape != (N, V, 3, 3)):
        raise ValueError("skinning: Input tensor shapes do not match the expected dimensions.")
    arclength = compute_arclength(strands)
    out = torch.zeros_like(strands)
    _core.skinning(strands, arclength, handle_strands, canonical_handles, R, out, beta)
    return out
