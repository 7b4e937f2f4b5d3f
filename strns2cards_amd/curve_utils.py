"""Strand utilities: frames, rotations, arclength, Gaussian smoothing,
rotation skinning, wisp alignment, tubes and outlier masks.

Mirror of the reference module pystrns2cards/curve_utils.py.  The Bishop
frame recurrence, smoothing and skinning run in HIP kernels; the per-vertex
frame algebra is batched torch; wisp alignment processes all wisps at once
(one host synchronization for the 3x3 root covariances) instead of a Python
loop with one SVD and several synchronizations per wisp.
"""
from typing import List, Tuple

import numpy as np
import torch

from . import _core

__all__ = ["minimal_rotation", "tangent_frame", "compute_bishop_frame", "compute_arclength",
           "smooth_strands", "skinning", "compute_aligned_strands", "compute_alignment_rotation",
           "strands_to_tube", "align_wisp", "align_wisps", "statistical_outlier_mask"]


def _unit(v: torch.Tensor) -> torch.Tensor:
    return v / v.norm(dim=-1, keepdim=True).clamp(min=1e-8)


def minimal_rotation(T_src: torch.Tensor, T_dst: torch.Tensor) -> torch.Tensor:
    """(*, 3, 3) rotation taking direction T_src to T_dst about their common
    normal (Rodrigues), as curve_utils.py:9-74: the identity when the
    directions agree (cos > 1 - 1e-6), and the fixed matrix diag(-1, -1, 1)
    when they are opposite (cos < -1 + 1e-6)."""
    a = _unit(T_src.float())
    b = _unit(T_dst.float())
    a, b = torch.broadcast_tensors(a, b)
    axis = torch.cross(a, b, dim=-1)
    s = axis.norm(dim=-1)
    c = (a * b).sum(-1).clamp(-1.0, 1.0)
    k = axis / s.clamp(min=1e-8)[..., None]
    zero = torch.zeros_like(k[..., 0])
    K = torch.stack([torch.stack([zero, -k[..., 2], k[..., 1]], -1),
                     torch.stack([k[..., 2], zero, -k[..., 0]], -1),
                     torch.stack([-k[..., 1], k[..., 0], zero], -1)], -2)
    eye = torch.eye(3, device=a.device, dtype=a.dtype).expand(K.shape)
    R = eye + s[..., None, None] * K + (1.0 - c)[..., None, None] * (K @ K)
    flip = torch.diag(torch.tensor([-1.0, -1.0, 1.0], device=a.device, dtype=a.dtype))
    R = torch.where((c > 1 - 1e-6)[..., None, None], eye, R)
    R = torch.where((c < -1 + 1e-6)[..., None, None], flip.expand(K.shape), R)
    return R


def tangent_frame(tangent: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """(normal, binormal) completing the unit tangent to a right-handed
    orthonormal basis without a branch on the direction (the construction
    of Duff et al., 2017, used by curve_utils.py:76-102)."""
    t = _unit(tangent)
    x, y, z = t.unbind(-1)
    sg = torch.where(z < 0, -torch.ones_like(z), torch.ones_like(z))
    a = -1.0 / (sg + z)
    b = x * y * a
    normal = torch.stack([1.0 + sg * x * x * a, sg * b, -sg * x], -1)
    binormal = torch.stack([b, sg + y * y * a, -y], -1)
    return normal, binormal


def compute_bishop_frame(strands: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(T, N, B) rotation-minimizing frames of (N, S, 3) strands by double
    reflection (curve_utils.py:104-165), computed by one HIP thread per
    strand.  Strands on the host are processed on the current GPU and the
    frames returned on the host in the input dtype."""
    if strands.ndim != 3 or strands.shape[-1] != 3:
        raise ValueError("compute_bishop_frame: strands must have shape (N, S, 3)")
    dev, dt = strands.device, strands.dtype
    x = strands.to(device="cuda" if not strands.is_cuda else dev, dtype=torch.float32).contiguous()
    T = torch.empty_like(x)
    Nr = torch.empty_like(x)
    Bn = torch.empty_like(x)
    _core.bishop_frame(x, T, Nr, Bn)
    if strands.is_cuda and dt == torch.float32:
        return T, Nr, Bn
    return T.to(dev, dt), Nr.to(dev, dt), Bn.to(dev, dt)


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


def compute_aligned_strands(strands: torch.Tensor, roots: torch.Tensor,
                            T_dst: torch.Tensor) -> torch.Tensor:
    """Straight strands along T_dst (1, 3) from ``roots`` (N, 3), each vertex
    at its arclength along the input strand (curve_utils.py:242-270)."""
    if strands.ndim != 3 or strands.shape[-1] != 3:
        raise ValueError("strands must have shape (N, V, 3)")
    if roots.ndim != 2 or roots.shape[-1] != 3:
        raise ValueError("roots must have shape (N, 3)")
    if T_dst.ndim != 2 or T_dst.shape != (1, 3):
        raise ValueError("T_dst must have shape (1, 3)")
    s = compute_arclength(strands)
    return roots[:, None, :] + s[:, :, None] * T_dst[None]


def compute_alignment_rotation(strands: torch.Tensor, T_dst: torch.Tensor) -> torch.Tensor:
    """(N, V, 3, 3) rotations taking each vertex's Bishop frame to the frame
    (T_dst, R0 N_root, R0 B_root), where R0 is the minimal rotation of the
    root tangent onto T_dst (curve_utils.py:272-308)."""
    if strands.ndim != 3 or strands.shape[-1] != 3:
        raise ValueError("strands must have shape (N, V, 3)")
    if T_dst.ndim != 2 or T_dst.shape != (1, 3):
        raise ValueError("T_dst must have shape (1, 3)")
    T, Nr, Bn = compute_bishop_frame(strands)
    td = T_dst.to(T.dtype).expand(T.shape[0], 3)
    R0 = minimal_rotation(T[:, 0], td)
    n_root = (R0 @ Nr[:, 0, :, None])[..., 0]
    b_root = (R0 @ Bn[:, 0, :, None])[..., 0]
    src = torch.stack([T, Nr, Bn], -1)  # columns T, N, B
    dst = torch.stack([td[:, None].expand_as(T), n_root[:, None].expand_as(T),
                       b_root[:, None].expand_as(T)], -1)
    return dst @ src.transpose(-1, -2)


def strands_to_tube(strands: np.ndarray, radius=1.0, n_circle=8, attr=None):
    """Open tubes of ``n_circle``-gon cross-sections around (N, S, 3) strands,
    oriented by their Bishop frames (curve_utils.py:310-372).  Returns
    (verts (N*S*n_circle, 3), faces (M, 3) int32[, attr expanded per
    vertex]); vertex (strand n, point i, corner k) is at
    (n*S + i)*n_circle + k."""
    x = np.asarray(strands)
    assert x.ndim == 3 and x.shape[2] == 3, f"Expected (N, S, 3), got {x.shape}"
    N, S, _ = x.shape
    _, Nr, Bn = compute_bishop_frame(torch.from_numpy(np.ascontiguousarray(x)))
    Nr, Bn = Nr.numpy(), Bn.numpy()
    ang = np.linspace(0.0, 2.0 * np.pi, n_circle, endpoint=False)
    ring = (np.cos(ang)[None, None, :, None] * Nr[:, :, None, :] +
            np.sin(ang)[None, None, :, None] * Bn[:, :, None, :])  # (N, S, k, 3)
    verts = (x[:, :, None, :] + radius * ring).reshape(N * S * n_circle, 3)
    cur = np.arange(max(S - 1, 0) * n_circle).reshape(-1, n_circle)
    nxt_k = cur - cur % n_circle + (cur + 1) % n_circle
    up, up_k = cur + n_circle, nxt_k + n_circle
    one = np.stack([np.stack([cur, up, up_k], -1), np.stack([cur, up_k, nxt_k], -1)], 0)
    one = one.reshape(-1, 3)
    faces = (one[None] + (np.arange(N) * S * n_circle)[:, None, None]).reshape(-1, 3)
    faces = faces.astype(np.int32)
    if attr is None:
        return verts, faces
    attr = np.asarray(attr)
    assert attr.shape[:2] == (N, S), f"Expected attr shape (N, S, C), got {attr.shape}"
    return verts, faces, np.repeat(attr[:, :, None, :], n_circle, axis=2).reshape(
        N * S * n_circle, attr.shape[-1])


def _sym_axes(cov: np.ndarray) -> np.ndarray:
    """Right singular vectors of the centred root matrices, from their 3x3
    Gram matrices: rows sorted by decreasing singular value."""
    w, v = np.linalg.eigh(cov)
    return np.swapaxes(v[..., ::-1], -1, -2)


def _root_gram(roots: torch.Tensor, seg: torch.Tensor, n_seg: int) -> np.ndarray:
    g = (roots[:, :, None] * roots[:, None, :]).reshape(-1, 9).double()
    out = torch.zeros((n_seg, 9), dtype=torch.float64, device=roots.device)
    out.index_add_(0, seg, g)
    return out.reshape(n_seg, 3, 3).cpu().numpy()


def align_wisp(strands: torch.Tensor, beta: float) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Aligns one wisp to the canonical frame (curve_utils.py:374-478):
    centre the roots, rotate the longest smoothed strand's root frame (the
    tangents of its first two segments with the roots' principal axis
    projected out) onto T=(0,-1,0), N=(0,0,-1), B=(1,0,0), then straighten
    by skinning the strands to their handles' arclength-parametrized lines.
    Returns (aligned, handles, index of the longest strand)."""
    if not strands.is_cuda:
        raise ValueError("Input strands must reside on CUDA.")
    if strands.ndim != 3 or strands.shape[-1] != 3:
        raise ValueError("Input tensor must have shape (N, V, 3)")
    n, v, _ = strands.shape
    dt, dev = strands.dtype, strands.device
    x = strands - strands[:, 0, :].mean(0, keepdim=True)
    roots = x[:, 0, :]
    gram = _root_gram(roots, torch.zeros(n, dtype=torch.long, device=dev), 1)
    axes = _sym_axes(gram)[0]
    # all roots equal (one strand): the reference's SVD of a zero matrix
    # returns the identity, so its first axis is x; eigh's order would give z
    rank0 = not np.any(gram)
    pa = axes[0] if np.all(np.isfinite(axes[0])) and not rank0 else np.array([1.0, 0.0, 0.0])
    pa = torch.as_tensor(pa, dtype=dt, device=dev)
    lengths = (x[:, 1:] - x[:, :-1]).norm(dim=2).sum(1)
    longest = torch.argmax(lengths)
    handles = smooth_strands(x, beta=beta)
    g = handles[longest]
    t0 = g[1] - g[0]
    t1 = g[2] - g[1]
    t0 = t0 - (t0 @ pa) * pa
    t1 = t1 - (t1 @ pa) * pa
    t0 = t0 / t0.norm()
    t1 = t1 / t1.norm()
    b0 = torch.cross(t0, t1, dim=-1)
    b0 = b0 / b0.norm()
    n0 = torch.cross(b0, t0, dim=-1)
    n0 = n0 / n0.norm()
    if not bool(torch.isfinite(n0).all() & torch.isfinite(b0).all()):
        n0, b0 = tangent_frame(t0[None])
        n0, b0 = n0[0], b0[0]
    if not bool(torch.isfinite(t0).all() & torch.isfinite(b0).all() & torch.isfinite(n0).all()):
        raise AssertionError("align_wisp: degenerate guide strand")
    Td = torch.tensor([0.0, -1.0, 0.0], dtype=dt, device=dev)
    Nd = torch.tensor([0.0, 0.0, -1.0], dtype=dt, device=dev)
    Bd = torch.tensor([1.0, 0.0, 0.0], dtype=dt, device=dev)
    Rg = torch.stack([Td, Nd, Bd], -1) @ torch.stack([t0, n0, b0], -1).T
    x = (x.reshape(-1, 3) @ Rg.T).reshape(n, v, 3)
    handles = (handles.reshape(-1, 3) @ Rg.T).reshape(n, v, 3)
    canon = compute_aligned_strands(handles, x[:, 0, :], Td[None])
    R = compute_alignment_rotation(handles, Td[None])
    return skinning(x, handles, canon, R, beta=beta), handles, longest


def align_wisps(wisps: List[np.ndarray], beta: float) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Aligns every wisp (list of (N_i, V, 3) arrays) to the canonical frame
    T=(0,1,0), N=(0,0,-1), B=(-1,0,0) (curve_utils.py:481-607): per wisp,
    N_src is the roots' least principal axis signed towards the longest
    strand's root->second-vertex direction, T_src is the root->tip direction
    of that strand with N_src projected out (the second principal axis if
    that degenerates), B_src = T_src x N_src; then all wisps are
    straightened by one skinning launch.  Wisps with fewer than 3 strands
    use the coordinate axes as principal axes.  Returns (aligned, smoothed
    handles before rotation, global index of each wisp's longest strand)."""
    dev = torch.device("cuda")
    sizes = [int(np.asarray(w).shape[0]) for w in wisps]
    x = torch.cat([torch.from_numpy(np.asarray(w)) for w in wisps], 0).to(device=dev,
                                                                           dtype=torch.float32)
    if x.ndim != 3 or x.shape[-1] != 3:
        raise ValueError("Input tensor must have shape (N, V, 3)")
    n, v, _ = x.shape
    nw = len(sizes)
    seg = torch.repeat_interleave(torch.arange(nw, device=dev), torch.tensor(sizes, device=dev))
    cnt = torch.tensor(sizes, device=dev, dtype=torch.float32)
    lengths = (x[:, 1:] - x[:, :-1]).norm(dim=2).sum(1)
    handles0 = smooth_strands(x, beta=beta)
    centre = torch.zeros((nw, 3), device=dev).index_add_(0, seg, x[:, 0, :]) / cnt[:, None]
    xc = x - centre[seg][:, None, :]
    hc = handles0 - centre[seg][:, None, :]
    # longest strand per wisp (first index on ties, as argmax)
    best = torch.full((nw,), -1.0, device=dev).scatter_reduce(0, seg, lengths, "amax")
    idx = torch.arange(n, device=dev)
    cand = torch.where(lengths == best[seg], idx, torch.full_like(idx, n))
    longest = torch.full((nw,), n, device=dev, dtype=torch.long).scatter_reduce(0, seg, cand, "amin")
    # principal axes from the 3x3 Gram matrices of the centred roots
    gram = _root_gram(xc[:, 0, :], seg, nw)
    axes = _sym_axes(gram)
    ok = (np.asarray(sizes) >= 3) & np.all(np.isfinite(axes), axis=(1, 2))
    axes = np.where(ok[:, None, None], axes, np.eye(3)[None])
    axes = torch.as_tensor(axes, dtype=torch.float32, device=dev)
    least, second = axes[:, 2], axes[:, 1]
    guide = xc[longest]  # (nw, V, 3)
    view = _unit(guide[:, 0] - guide[:, 1])
    down = _unit(guide[:, -1] - guide[:, 0])
    sgn = torch.sign((least * view).sum(-1))
    sgn = torch.where(sgn == 0, torch.ones_like(sgn), sgn)
    Ns = least * sgn[:, None]
    Ts = down - (down * Ns).sum(-1, keepdim=True) * Ns
    Ts = Ts / Ts.norm(dim=-1, keepdim=True)
    Ts = torch.where(torch.isfinite(Ts).all(-1, keepdim=True), Ts, second)
    Bs = torch.cross(Ts, Ns, dim=-1)
    Td = torch.tensor([0.0, 1.0, 0.0], device=dev)
    Nd = torch.tensor([0.0, 0.0, -1.0], device=dev)
    Bd = torch.tensor([-1.0, 0.0, 0.0], device=dev)
    Rdst = torch.stack([Td, Nd, Bd], -1)
    Rg = Rdst[None] @ torch.stack([Ts, Ns, Bs], -1).transpose(-1, -2)  # (nw, 3, 3)
    Rs = Rg[seg]
    xr = (xc @ Rs.transpose(-1, -2))
    hr = (hc @ Rs.transpose(-1, -2))
    canon = compute_aligned_strands(hr, xr[:, 0, :], Td[None])
    R = compute_alignment_rotation(hr, Td[None])
    return skinning(xr, hr, canon, R, beta=beta), handles0, longest


def statistical_outlier_mask(points: np.ndarray, n_neighbors: int = 5,
                             ratio: float = 2.0) -> np.ndarray:
    """(N,) bool: True where a point's mean distance to its k nearest
    neighbours (k = min(N-1, n_neighbors), itself excluded) is below
    ratio x the mean of those means (curve_utils.py:610-644).  Exact k-NN
    by blocked brute force in float64 on the GPU."""
    p = np.asarray(points)
    N = p.shape[0]
    if N <= 1:
        return np.ones(N, dtype=np.bool_)
    k = min(N - 1, n_neighbors)
    x = torch.from_numpy(np.ascontiguousarray(p, dtype=np.float64)).cuda()
    means = torch.empty(N, dtype=torch.float64, device=x.device)
    rows = max(1, (1 << 24) // max(N, 1))
    for i in range(0, N, rows):
        d = torch.cdist(x[i:i + rows], x, compute_mode="donot_use_mm_for_euclid_dist")
        near = torch.topk(d, k + 1, dim=1, largest=False).values  # sorted ascending
        means[i:i + rows] = near[:, 1:].mean(1)
    thr = ratio * means.mean()
    return (means < thr).cpu().numpy()
