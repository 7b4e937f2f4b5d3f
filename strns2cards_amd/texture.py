"""Software-MSAA texture patches of wisps (render_texture,
texture_utils.py:106-288).

The reference renders each wisp's tube mesh once per sample of an 8 x 8
jitter grid spanning two pixels, 64 separate ``rasterize`` +
``interpolate_triangle_attributes`` calls with host synchronizations, and
accumulates with torch ops.  ``render_patch`` makes the 64 jittered vertex
sets in one kernel, depth-tests them as 64 views in one kernel that keeps
only the winning (depth, triangle) key per sample and pixel (or, with
``fused=False``, rasterizes them as one batched ``rasterize``), and
resolves every pixel over its samples in one kernel:

* rgb: the mean of the interpolated attribute over the covered samples,
  (0, 1, 0) where none is covered, mapped by 0.5 (rgb + 1);
* alpha: the Mitchell-Netravali (B = C = 1/3) weighted coverage, blurred
  by a normalized 3-tap Gaussian (sigma 1.5, reflect borders) as kornia's
  ``gaussian_blur2d`` with kernel (3, 3) does.

``render_texture`` runs the reference's per-wisp loop on top of it and
pastes each patch (flipped vertically, as bytes) into the atlas on the GPU.

``patch_masks`` renders the wisp masks of ``render_patches_reduced``
(texture_utils.py:290-411), all wisps in one launch, with the same
depth-key idea for DDA fragments.
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch

from . import _core
from .curve_utils import strands_to_tube
from .rasterize import rasterize


def mitchell_netravali(x, B: float = 1.0 / 3, C: float = 1.0 / 3):
    """The cubic reconstruction filter of texture_utils.py:189-201."""
    ax = np.abs(np.asarray(x, np.float64))
    near = ((12 - 9 * B - 6 * C) * ax ** 3 + (-18 + 12 * B + 6 * C) * ax ** 2 + (6 - 2 * B))
    far = ((-B - 6 * C) * ax ** 3 + (6 * B + 30 * C) * ax ** 2 + (-12 * B - 48 * C) * ax
           + (8 * B + 24 * C))
    return (near * (ax < 1) + far * ((ax >= 1) & (ax < 2))) / 6.0


def msaa_samples(grid_res: int = 8, span: float = 2.0):
    """Sample offsets in pixels (y-major, x fastest) and their normalized
    filter weights, as texture_utils.py:203-215."""
    g = span * ((np.arange(grid_res) + 0.5) / grid_res - 0.5)
    offsets = [(float(dx), float(dy)) for dy in g for dx in g]
    w = np.array([mitchell_netravali(dx) * mitchell_netravali(dy) for dx, dy in offsets])
    return offsets, (w / w.sum()).tolist()


def gaussian_kernel3(sigma: float = 1.5) -> list[float]:
    """Normalized float32 taps at -1, 0, 1 of exp(-x^2 / (2 sigma^2)), the
    1-D kernel kornia's gaussian_blur2d builds for a kernel size of 3."""
    x = np.arange(3, dtype=np.float32) - np.float32(1)
    s = np.float32(sigma)
    g = np.exp(-(x * x) / (np.float32(2) * s * s)).astype(np.float32)
    return (g / g.sum(dtype=np.float32)).astype(np.float32).tolist()


def render_patch(pos: torch.Tensor, tri: torch.Tensor, attr: torch.Tensor,
                 patch_res: tuple[int, int], grid_res: int = 8, span: float = 2.0,
                 fused: bool = True):
    """(rgb (H, W, 3), alpha (H, W)) float32 device tensors of one wisp's
    NDC mesh (pos (V, 3), tri (T, 3) int32, attr (V, 3)): the patch of
    texture_utils.py:203-260 before its vertical flip.  `fused` (default)
    keeps only a depth key per (sample, pixel) instead of the batched
    rasterize's fragment lists; both give the same patch."""
    H, W = patch_res
    dev = pos.device
    offsets, weights = msaa_samples(grid_res, span)
    # the reference adds Python floats to a float32 tensor: float32 offsets
    dx = torch.tensor([o[0] * 2.0 / W for o in offsets], dtype=torch.float32, device=dev)
    dy = torch.tensor([o[1] * 2.0 / H for o in offsets], dtype=torch.float32, device=dev)
    w = torch.tensor(weights, dtype=torch.float32, device=dev)
    S = len(offsets)
    pos = pos.contiguous().float()
    attr = attr.contiguous().float()
    tri = tri.contiguous()
    jit = torch.empty((S, pos.shape[0], 4), dtype=torch.float32, device=dev)
    _core.msaa_jitter(pos, dx, dy, jit)
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    alpha = torch.empty((H, W), dtype=torch.float32, device=dev)
    if fused:
        keys = torch.empty((S, H, W), dtype=torch.int64, device=dev)
        _core.msaa_patch(jit, tri, attr, w, H, W, keys, rgb, alpha)
    else:
        rast = rasterize((H, W), jit, tri)
        _core.msaa_resolve(rast, attr, tri, w, rgb, alpha)
    tmp = torch.empty_like(alpha)
    blurred = torch.empty_like(alpha)
    _core.blur3x3_reflect(gaussian_kernel3(1.5), alpha, tmp, blurred)
    return rgb, blurred


def atlas_rect(uv_rect, resolution: int, patch_res: tuple[int, int]):
    """Atlas rectangle of a wisp and the crop of its patch, as
    texture_utils.py:264-276: (crop_x0, crop_y0, x0, y0, x1, y1)."""
    u0, v0, u1, v1 = uv_rect
    x0, y0 = int(np.floor(u0 * resolution)), int(np.floor(v0 * resolution))
    x1, y1 = int(np.floor(u1 * resolution)), int(np.floor(v1 * resolution))
    crop_x0 = int((patch_res[1] - (x1 - x0)) / 2)
    return crop_x0, 0, x0, y0, x1, y1


def packing_frame(packed_wisps):
    """(points, point offsets, packed_bbox_size, to_ndc_scale) of the packed
    wisps (texture_utils.py:120-141): the xy extent of all points, and 2
    over the largest per-wisp extent (z scaled by 0.2 more).  Per-wisp
    extents come from one segmented min / max over the concatenated points
    (the same exact values as the reference's per-wisp reductions)."""
    pts = np.concatenate([np.asarray(w).reshape(-1, 3) for w in packed_wisps], axis=0)
    counts = np.array([np.asarray(w).reshape(-1, 3).shape[0] for w in packed_wisps])
    if counts.min() < 1:
        raise ValueError("every packed wisp needs at least one point")
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    packed_bbox_size = pts[:, :2].max(axis=0) - pts[:, :2].min(axis=0)
    ext = np.maximum.reduceat(pts, offs, axis=0) - np.minimum.reduceat(pts, offs, axis=0)
    bbox_size = np.maximum(np.zeros(3), ext.max(axis=0))
    to_ndc_scale = 2.0 * np.ones(3) / bbox_size
    to_ndc_scale[2] *= 0.2
    return pts, counts, packed_bbox_size, to_ndc_scale


def wisp_ndc_mesh(wisp: np.ndarray, uv_rect, packed_bbox_size, to_ndc_scale, tube_radius,
                  n_circle):
    """The tube mesh of one (N, S, 3) wisp with its tangents as attribute,
    placed in the patch's NDC frame (texture_utils.py:155-178; host numpy
    in the reference, kept so)."""
    tangents = np.zeros_like(wisp)
    deltas = np.diff(wisp, axis=1)
    tangents[:, :-1, :] += deltas
    tangents[:, 1:, :] += deltas
    tangents /= np.linalg.norm(tangents, axis=2, keepdims=True)
    verts, faces, attr = strands_to_tube(wisp, radius=tube_radius, n_circle=n_circle,
                                         attr=tangents)
    u0, v0, u1, v1 = uv_rect
    u0w, u1w = u0 * packed_bbox_size[0], u1 * packed_bbox_size[0]
    v0w = v0 * packed_bbox_size[1]
    verts[:, 0] -= 0.5 * (u0w + u1w)
    verts[:, 1] -= v0w
    verts *= to_ndc_scale
    verts[:, 1] -= 1.0
    return verts, faces, attr


def render_texture(packed_wisps, uv_coords, tube_radius, n_circle, resolution=1024,
                   output_dirname="./output", subsampling_ratio=None, device="cuda"):
    """texture_utils.py:106-288: the tangent (rgb) and alpha atlases of the
    packed wisps, (resolution, resolution, 3) and (resolution, resolution)
    uint8 numpy arrays.  Written as tangent/alpha .png/.tga when
    `output_dirname` is given and imageio is importable."""
    max_u = max(u1 - u0 for _, u0, _, u1, _ in uv_coords)
    max_v = max(v1 - v0 for _, _, v0, _, v1 in uv_coords)
    patch_res = (int(math.ceil(max_v * resolution)), int(math.ceil(max_u * resolution)))
    _, _, packed_bbox_size, to_ndc_scale = packing_frame(packed_wisps)
    dev = torch.device(device)
    atlas_rgb = torch.zeros((resolution, resolution, 3), dtype=torch.uint8, device=dev)
    atlas_alpha = torch.zeros((resolution, resolution), dtype=torch.uint8, device=dev)
    for wisp, (_, u0, v0, u1, v1) in zip(packed_wisps, uv_coords):
        if subsampling_ratio is not None and subsampling_ratio < 1.0:
            S = wisp.shape[0]
            keep = np.random.choice(S, size=max(1, int(S * subsampling_ratio)), replace=False)
            wisp = wisp[keep]
        verts, faces, attr = wisp_ndc_mesh(wisp, (u0, v0, u1, v1), packed_bbox_size,
                                           to_ndc_scale, tube_radius, n_circle)
        rgb, alpha = render_patch(torch.from_numpy(verts).to(dev, torch.float32),
                                  torch.from_numpy(faces).to(dev, torch.int32),
                                  torch.from_numpy(np.ascontiguousarray(attr)).to(dev, torch.float32),
                                  patch_res)
        cx0, cy0, x0, y0, x1, y1 = atlas_rect((u0, v0, u1, v1), resolution, patch_res)
        _core.atlas_paste(rgb, alpha, cx0, cy0, x0, y0, x1, y1, atlas_rgb, atlas_alpha)
    rgb_np, alpha_np = atlas_rgb.cpu().numpy(), atlas_alpha.cpu().numpy()
    if output_dirname is not None:
        try:
            import imageio.v3 as iio
        except ImportError:
            iio = None
        if iio is not None:
            os.makedirs(output_dirname, exist_ok=True)
            for ext in ("png", "tga"):
                iio.imwrite(os.path.join(output_dirname, f"tangent.{ext}"), rgb_np)
                iio.imwrite(os.path.join(output_dirname, f"alpha.{ext}"), alpha_np)
    return rgb_np, alpha_np


def wisp_ndc_points(wisp: np.ndarray, uv_rect, packed_bbox_size, to_ndc_scale) -> np.ndarray:
    """(N*S, 3) float32 strand points of one wisp in its patch's NDC frame
    (texture_utils.py:334-343; host numpy in the reference, kept so)."""
    u0, v0, u1, v1 = uv_rect
    verts = np.array(wisp, copy=True).reshape(-1, 3)
    u0w, u1w = u0 * packed_bbox_size[0], u1 * packed_bbox_size[0]
    verts[:, 0] -= 0.5 * (u0w + u1w)
    verts[:, 1] -= v0 * packed_bbox_size[1]
    verts *= to_ndc_scale
    verts[:, 1] -= 1.0
    return verts.astype(np.float32)


def patch_masks_device(pos: torch.Tensor, rows, cols, patch_res: tuple[int, int]) -> torch.Tensor:
    """(num_wisps, H, W) uint8 device masks of wisps whose clip-space points
    (float4, w = 1) lie back to back in `pos` (wisp g: rows[g] strands of
    cols[g] points): one launch for all wisps."""
    H, W = patch_res
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    counts = rows * cols
    voff = np.zeros_like(counts)
    if counts.size:
        voff[1:] = np.cumsum(counts)[:-1]
    dev = pos.device
    G = int(rows.size)
    keys = torch.empty((G, H, W), dtype=torch.int64, device=dev)
    out = torch.empty((G, H, W), dtype=torch.uint8, device=dev)
    max_edges = int((rows * np.maximum(cols - 1, 0)).max()) if G else 0
    _core.patch_masks(torch.from_numpy(rows.astype(np.int32)).to(dev),
                      torch.from_numpy(cols.astype(np.int32)).to(dev),
                      torch.from_numpy(voff).to(dev), max_edges, pos.contiguous(), keys, out)
    return out


def patch_masks(packed_wisps, uv_coords, resolution, device="cuda") -> np.ndarray:
    """The per-wisp masks of render_patches_reduced (texture_utils.py:
    304-358) as one (num_wisps, H, W) uint8 array: every wisp's strands
    DDA-rasterized, depth-tested and interpolated (a constant 1) into the
    shared patch resolution, flipped and quantized, all wisps in one
    launch."""
    max_u = max(u1 - u0 for _, u0, _, u1, _ in uv_coords)
    max_v = max(v1 - v0 for _, _, v0, _, v1 in uv_coords)
    H, W = int(math.ceil(max_v * resolution)), int(math.ceil(max_u * resolution))
    _, _, packed_bbox_size, to_ndc_scale = packing_frame(packed_wisps)
    # per wisp, so that the shifts keep the reference's scalar types (numpy
    # float64 uv values make the float32 subtraction round once from double)
    allp = np.concatenate([wisp_ndc_points(w, r[1:5], packed_bbox_size, to_ndc_scale)
                           for w, r in zip(packed_wisps, uv_coords)])
    dev = torch.device(device)
    pos = torch.ones((allp.shape[0], 4), dtype=torch.float32, device=dev)
    pos[:, :3] = torch.from_numpy(allp).to(dev)
    rows = [w.shape[0] for w in packed_wisps]
    cols = [w.shape[1] for w in packed_wisps]
    return patch_masks_device(pos, rows, cols, (H, W)).cpu().numpy()


def render_patches_reduced(packed_wisps, uv_coords, resolution, n_clusters=20,
                           output_dirname="./output", dump_patches=False):
    """texture_utils.py:290-411: the wisp masks (on the GPU, `patch_masks`),
    then k-means over them (scikit-learn, as the reference) and the member
    of each cluster closest to its binarized centre.  Returns
    (centroid_indices, labels); the mask images are written when imageio is
    importable."""
    from sklearn.cluster import KMeans

    masks = patch_masks(packed_wisps, uv_coords, resolution)
    mask_all = masks.reshape(len(masks), -1)
    kmeans = KMeans(n_clusters=n_clusters, random_state=42)
    labels = kmeans.fit_predict(mask_all)
    centroid_indices = []
    for k in range(n_clusters):
        member = labels == k
        pts = mask_all[member]
        if len(pts) == 0:
            continue
        center = kmeans.cluster_centers_[k].copy()
        center[center > 0] = 1.0
        dist = np.sum(np.minimum(pts, center), axis=1)
        centroid_indices.append(np.where(member)[0][np.argmin(dist)])
    try:
        import imageio.v3 as iio
    except ImportError:
        iio = None
    if iio is not None and output_dirname is not None:
        if dump_patches:
            os.makedirs(os.path.join(output_dirname, "tmp"), exist_ok=True)
            for idx, m in enumerate(masks):
                iio.imwrite(os.path.join(output_dirname, "tmp", f"mask_{idx}.png"), m)
        os.makedirs(os.path.join(output_dirname, "centroids"), exist_ok=True)
        for i, idx in enumerate(centroid_indices):
            iio.imwrite(os.path.join(output_dirname, "centroids", f"centroid_{i}.png"), masks[idx])
    return centroid_indices, labels

