"""Card mesh inputs on the GPU.

The scripts build the strips and meshes the fitting consumes on the host
with numpy: ``strand_to_strip_mesh`` (2_texture.py:17-96),
``wisp_to_grid_mesh`` / ``combine_wisps_with_cluster_labels``
(texture_utils.py:525-655) and ``build_clustered_strand_edges``
(3_fitting.py:17-60).  The same functions here keep their names, arguments
and return layouts (numpy arrays) and compute on the GPU; the ``*_device``
forms take and return device tensors for many strands / grids at once, so
a job's inputs never leave HBM.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _core


def _ragged(strands, device):
    """(points (P,3) f32 tensor, counts (G,) i32, offsets (G,) i64) of a
    list of (n_i, 3) arrays."""
    arrs = [np.asarray(s, np.float32).reshape(-1, 3) for s in strands]
    counts = np.array([a.shape[0] for a in arrs], np.int64)
    offsets = np.zeros_like(counts)
    if counts.size:
        offsets[1:] = np.cumsum(counts)[:-1]
    pts = np.concatenate(arrs) if arrs else np.zeros((0, 3), np.float32)
    dev = torch.device(device)
    return (torch.from_numpy(pts).to(dev), torch.from_numpy(counts.astype(np.int32)).to(dev),
            torch.from_numpy(offsets).to(dev), counts)


def strands_to_strips_device(points: torch.Tensor, counts: torch.Tensor, offsets: torch.Tensor,
                             res_w: int = 2, width: float = 0.5,
                             tangent_rotations_deg=(0.0,)) -> torch.Tensor:
    """Strips of many strands (strand s: counts[s] >= 2 points of `points`
    from offsets[s]) as one (V, 3) float32 vertex array, in the order
    2_texture.py:304-312 collects them (strand-major, then rotation; each
    strip res_w rows by counts[s] columns).  Its grid layout is rows =
    res_w and cols = counts repeated per rotation."""
    rot = [float(a) for a in tangent_rotations_deg]
    if counts.numel() and int(counts.min()) < 2:
        raise ValueError("strands_to_strips: every strand needs at least 2 points")
    out = torch.empty((points.shape[0] * res_w * len(rot), 3), dtype=torch.float32,
                      device=points.device)
    _core.strands_to_strips(counts, offsets, points.contiguous(), int(res_w), float(width), rot,
                            out)
    return out


def strand_to_strip_mesh(strand: np.ndarray, res_w: int = 2, width: float = 0.5,
                         tangent_rotations_deg: list = [0.0]) -> list[np.ndarray]:  # noqa: B006
    """2_texture.py:17-96 for one strand: a list of (res_w, S, 3) float32
    strips, one per tangent rotation."""
    if res_w < 2:
        raise AssertionError("res_w must be at least 2")
    s = np.asarray(strand, np.float32)
    if s.ndim != 2 or s.shape[1] != 3 or s.shape[0] < 2:
        raise ValueError("strand_to_strip_mesh: strand must be (S >= 2, 3)")
    pts, counts, offsets, _ = _ragged([s], "cuda")
    out = strands_to_strips_device(pts, counts, offsets, res_w, width, tangent_rotations_deg)
    out = out.cpu().numpy().reshape(len(tangent_rotations_deg), res_w, s.shape[0], 3)
    return [out[r] for r in range(out.shape[0])]


def grid_layout(rows, cols):
    """Offsets of concatenated grid meshes: vertices, quads, h_edges,
    w_edges (int64 numpy, the prefix of each grid's counts) and totals."""
    rows = np.asarray(rows, np.int64).reshape(-1)
    cols = np.asarray(cols, np.int64).reshape(-1)
    if rows.size and (rows.min() < 1 or cols.min() < 1):
        raise ValueError("grid_layout: every grid needs at least one row and one column")
    counts = {"v": rows * cols, "q": (rows - 1) * (cols - 1), "h": rows * (cols - 1),
              "w": (rows - 1) * cols}
    offs, tot = {}, {}
    for k, c in counts.items():
        o = np.zeros_like(c)
        if c.size:
            o[1:] = np.cumsum(c)[:-1]
        offs[k], tot[k] = o, int(c.sum())
    if tot["v"] >= 2**31 or 2 * tot["q"] >= 2**31:
        raise ValueError("grid_layout: more than 2^31 - 1 vertices or faces")
    return offs, tot


def grid_mesh_topology_device(rows, cols, rects=None, grid2cluster=None, margin_top: float = 1e-4,
                              device="cuda") -> dict:
    """faces, quads, h_edges, w_edges (int32), and with `grid2cluster` the
    face clusters (int64), with `rects` ([u0, v0, u1, v1] per grid) the uvs
    (float64), of the concatenated grid meshes (texture_utils.py:525-655),
    as device tensors."""
    offs, tot = grid_layout(rows, cols)
    dev = torch.device(device)
    i32 = lambda a: torch.from_numpy(np.asarray(a, np.int64).astype(np.int32)).to(dev)  # noqa: E731
    out = {"faces": torch.empty((2 * tot["q"], 3), dtype=torch.int32, device=dev),
           "quads": torch.empty((tot["q"], 4), dtype=torch.int32, device=dev),
           "h_edges": torch.empty((tot["h"], 2), dtype=torch.int32, device=dev),
           "w_edges": torch.empty((tot["w"], 2), dtype=torch.int32, device=dev)}
    g2c = f2c = r = uvs = None
    if grid2cluster is not None:
        g2c = torch.as_tensor(np.asarray(grid2cluster, np.int64), device=dev)
        f2c = out["face2cluster"] = torch.empty(2 * tot["q"], dtype=torch.int64, device=dev)
    if rects is not None:
        r = torch.as_tensor(np.asarray(rects, np.float64).reshape(-1, 4), device=dev)
        uvs = out["uvs"] = torch.empty((tot["v"], 2), dtype=torch.float64, device=dev)
    _core.grid_mesh_topology(i32(rows), i32(cols), i32(offs["v"]), i32(offs["q"]), i32(offs["h"]),
                             i32(offs["w"]), g2c, r, float(margin_top), out["faces"],
                             out["quads"], out["h_edges"], out["w_edges"], f2c, uvs)
    return out


def wisp_to_grid_mesh(wisp: np.ndarray, uv_rect=None, margin_top: float = 1e-4):
    """texture_utils.py:525-599: (verts, faces, [uvs,] quads, h_edges,
    w_edges) of one (S, N, 3) wisp, numpy."""
    S, N, _ = wisp.shape
    m = grid_mesh_topology_device([S], [N], None if uv_rect is None else [uv_rect],
                                  margin_top=margin_top)
    verts = np.asarray(wisp).reshape(-1, 3)
    np_ = {k: v.cpu().numpy() for k, v in m.items()}
    if uv_rect is None:
        return verts, np_["faces"], np_["quads"], np_["h_edges"], np_["w_edges"]
    return verts, np_["faces"], np_["uvs"], np_["quads"], np_["h_edges"], np_["w_edges"]


def combine_wisps_with_cluster_labels(wisps, layout, wisp2cluster):
    """texture_utils.py:597-655: (vertices, faces, uvs, face2cluster, quads,
    h_edges, w_edges) of the wisps' grid meshes concatenated, numpy.
    `layout` holds (idx, u0, v0, u1, v1) per wisp."""
    shapes = np.array([w.shape[:2] for w in wisps], np.int64).reshape(-1, 2)
    rects = [tuple(entry[1:5]) for entry in layout]
    m = grid_mesh_topology_device(shapes[:, 0], shapes[:, 1], rects, list(wisp2cluster))
    verts = np.concatenate([np.asarray(w).reshape(-1, 3) for w in wisps], axis=0)
    c = {k: v.cpu().numpy() for k, v in m.items()}
    return verts, c["faces"], c["uvs"], c["face2cluster"], c["quads"], c["h_edges"], c["w_edges"]


def strand_edges_device(counts: torch.Tensor, offsets: torch.Tensor, cluster_ids: torch.Tensor,
                        num_edges: int):
    """(edges (E,2) int32, edge2cluster (E,) int32) between consecutive
    points of each strand (counts >= 1), as device tensors."""
    dev = counts.device
    edges = torch.empty((num_edges, 2), dtype=torch.int32, device=dev)
    e2c = torch.empty(num_edges, dtype=torch.int32, device=dev)
    _core.strand_edges(counts, offsets, cluster_ids.to(torch.int32), edges, e2c)
    return edges, e2c


def build_clustered_strand_edges(strands, strand_cluster_ids):
    """3_fitting.py:17-60: (points, edges (E,2) int32, edge_cluster_ids
    (E,) int32, num_clusters), numpy."""
    if len(strands) != len(strand_cluster_ids):
        raise AssertionError("Mismatch between strands and cluster IDs")
    pts, counts, offsets, counts_np = _ragged(strands, "cuda")
    if counts_np.size and counts_np.min() < 1:
        raise ValueError("build_clustered_strand_edges: every strand needs a point")
    ids = torch.as_tensor(np.asarray(strand_cluster_ids, np.int32), device=pts.device)
    edges, e2c = strand_edges_device(counts, offsets, ids, int((counts_np - 1).sum()))
    num_clusters = max([int(c) + 1 for c in strand_cluster_ids], default=0)
    points = np.concatenate([np.asarray(s) for s in strands], axis=0)
    return points, edges.cpu().numpy(), e2c.cpu().numpy(), num_clusters
