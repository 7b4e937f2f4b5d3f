"""Seeded synthetic workloads for the hot path (no dataset is available).

* strands: roots uniform on the upper (y >= 0) hemisphere of radius 0.1; the
  tangent starts at the surface normal and, per step, is rotated by an angle
  ~ N(0, 0.15 rad) about a uniformly random axis perpendicular to it, biased
  by 0.02 * (0, -1, 0) (gravity) and renormalized;
* clusters: an 8 x 8 azimuth / elevation grid of the roots, or Lloyd k-means
  of the roots with a fixed seed and iteration count;
* cameras: orthographic views built like ``build_ortho_cameras``
  (3_fitting.py:62-152): bbox-centred, uniform XY scale 2(1 - margin) / max
  extent, Z mapped to [-0.5, 0.5]; the first ``B`` of its 13 directions;
* cards: one strip per cluster: the pointwise mean strand, resampled by
  arclength to Q + 1 points, swept along a rotation-minimizing frame
  (``strand_to_strip_mesh``, 2_texture.py:17-96, res_w = 2) and triangulated
  as ``wisp_to_grid_mesh`` (texture_utils.py:525-599); edges follow
  ``build_edges_from_triangles`` (mesh_utils.cpp:15-79) with a deterministic
  first-appearance order.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# ----------------------------------------------------------------- strands


def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.maximum(np.linalg.norm(v, axis=-1, keepdims=True), 1e-12)


def generate_strands(num_strands: int, num_vertices: int, step: float, seed: int,
                     rng: np.random.Generator | None = None,
                     root_dirs: np.ndarray | None = None) -> np.ndarray:
    """(N, n_v, 3) float32 polylines (see module docstring)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    if root_dirs is None:
        d = rng.normal(size=(num_strands, 3))
        d = _normalize(d)
        d[:, 1] = np.abs(d[:, 1])
    else:
        d = _normalize(np.asarray(root_dirs, np.float64))
    x = 0.1 * d
    t = d.copy()
    out = np.empty((num_strands, num_vertices, 3), np.float64)
    out[:, 0] = x
    g = np.array([0.0, -0.02, 0.0])
    for i in range(1, num_vertices):
        r = rng.normal(size=(num_strands, 3))
        axis = _normalize(np.cross(t, r))
        ang = rng.normal(0.0, 0.15, size=(num_strands, 1))
        c, s = np.cos(ang), np.sin(ang)
        # Rodrigues rotation of t about axis (axis is perpendicular to t).
        t = t * c + np.cross(axis, t) * s
        t = _normalize(t + g)
        x = x + step * t
        out[:, i] = x
    return out.astype(np.float32)


def grid_clusters(roots: np.ndarray, n_az: int = 8, n_el: int = 8) -> np.ndarray:
    """Cluster id az * n_el + el from the root's azimuth atan2(z, x) in
    [-pi, pi) and elevation asin(y / |r|) in [0, pi/2]."""
    r = np.linalg.norm(roots, axis=1)
    az = np.arctan2(roots[:, 2], roots[:, 0])
    el = np.arcsin(np.clip(roots[:, 1] / np.maximum(r, 1e-12), -1, 1))
    ia = np.clip(((az + np.pi) / (2 * np.pi) * n_az).astype(np.int64), 0, n_az - 1)
    ie = np.clip((el / (np.pi / 2) * n_el).astype(np.int64), 0, n_el - 1)
    return (ia * n_el + ie).astype(np.int32)


def kmeans_clusters(points: np.ndarray, k: int, seed: int, iters: int = 12) -> np.ndarray:
    """Lloyd's k-means (k distinct random points as seeds, fixed iteration
    count); empty clusters keep their previous centre.  Labels are then
    renumbered by first appearance so ids are dense in [0, k)."""
    rng = np.random.default_rng(seed)
    p = np.asarray(points, np.float64)
    c = p[rng.choice(p.shape[0], size=k, replace=False)].copy()
    pn = (p * p).sum(1)
    lab = np.zeros(p.shape[0], np.int64)
    for _ in range(iters):
        cn = (c * c).sum(1)
        for s in range(0, p.shape[0], 8192):
            d = pn[s:s + 8192, None] - 2.0 * p[s:s + 8192] @ c.T + cn[None]
            lab[s:s + 8192] = d.argmin(1)
        cnt = np.bincount(lab, minlength=k)
        for dim in range(3):
            sm = np.bincount(lab, weights=p[:, dim], minlength=k)
            c[:, dim] = np.where(cnt > 0, sm / np.maximum(cnt, 1), c[:, dim])
    _, first = np.unique(lab, return_index=True)
    order = np.argsort(first)
    remap = np.full(k, -1, np.int64)
    remap[np.unique(lab)[order]] = np.arange(order.size)
    return remap[lab].astype(np.int32)


# ----------------------------------------------------------------- cameras

_VIEW_DIRS = [  # 3_fitting.py:84-103, as multiples of the half extents
    (1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1),
    (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1),
]


def ortho_cameras(points: np.ndarray, num_views: int, margin: float = 0.3) -> np.ndarray:
    """(num_views, 4, 4) MVP matrices (float64)."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    lo, hi = p.min(0), p.max(0)
    center = (lo + hi) / 2
    half = (hi - lo) / 2
    mats = []
    for k in range(num_views):
        vp = np.array(_VIEW_DIRS[k], np.float64) * half
        direction = -vp / np.linalg.norm(vp)
        up = np.array([0.0, 1.0, 0.0])
        if abs(direction @ up) > 0.9:
            up = np.array([1.0, 0.0, 0.0])
        right = np.cross(up, direction)
        right /= np.linalg.norm(right)
        up = np.cross(direction, right)
        R = np.stack([right, up, direction], axis=1)
        view = np.eye(4)
        view[:3, :3] = R.T
        view[:3, 3] = -R.T @ center
        cam = (p - center) @ R
        cmin, cmax = cam.min(0), cam.max(0)
        ext = cmax - cmin
        sxy = 2.0 * (1.0 - margin) / max(ext[0], ext[1])
        sz = 1.0 / max(ext[2], 1e-12)
        proj = np.eye(4)
        proj[0, 0] = proj[1, 1] = sxy
        proj[2, 2] = sz
        proj[:2, 3] = -(cmin[:2] + cmax[:2]) / 2 * sxy
        proj[2, 3] = -(cmin[2] + cmax[2]) / 2 * sz
        mats.append(proj @ view)
    return np.stack(mats)


def project(points: np.ndarray, mvp: np.ndarray) -> np.ndarray:
    """(B, V, 4) float32 clip coordinates of (V, 3) points."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    ph = np.concatenate([p, np.ones((p.shape[0], 1))], 1)
    return np.einsum("vk,bjk->bvj", ph, mvp).astype(np.float32)


# ------------------------------------------------------------------- cards


def resample_by_arclength(curves: np.ndarray, num: int) -> np.ndarray:
    """(C, n, 3) -> (C, num, 3) evenly spaced in arclength."""
    c = np.asarray(curves, np.float64)
    seg = np.linalg.norm(np.diff(c, axis=1), axis=2)
    s = np.concatenate([np.zeros((c.shape[0], 1)), np.cumsum(seg, 1)], 1)
    out = np.empty((c.shape[0], num, 3))
    for i in range(c.shape[0]):
        tgt = np.linspace(0.0, s[i, -1], num)
        for d in range(3):
            out[i, :, d] = np.interp(tgt, s[i], c[i, :, d])
    return out


def strip_rows(curves: np.ndarray, width: float) -> np.ndarray:
    """(C, n, 3) centre curves -> (C, 2, n, 3) strip rows offset by -w/2 and
    +w/2 along a rotation-minimizing binormal (double reflection)."""
    c = np.asarray(curves, np.float64)
    C, n, _ = c.shape
    d = np.diff(c, axis=1)
    tan = np.zeros_like(c)
    tan[:, :-1] += d
    tan[:, 1:] += d
    tan = _normalize(tan)
    T0 = tan[:, 0]
    dT = tan[:, 1] - tan[:, 0]
    base = np.where((np.abs(T0 @ np.array([0.0, 1.0, 0.0])) > 0.9)[:, None],
                    np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0]))
    small = np.linalg.norm(dT, axis=1) < 1e-4
    N0 = np.where(small[:, None], base, -_normalize(dT))
    B0 = _normalize(np.cross(T0, N0))
    N0 = np.cross(B0, T0)
    N = np.zeros_like(c)
    Bn = np.zeros_like(c)
    N[:, 0] = N0
    Bn[:, 0] = np.cross(T0, N0)
    for i in range(1, n):
        v = c[:, i] - c[:, i - 1]
        vn = np.linalg.norm(v, axis=1, keepdims=True)
        v = v / np.maximum(vn, 1e-12)
        nL = N[:, i - 1] - 2 * (N[:, i - 1] * v).sum(1, keepdims=True) * v
        tL = tan[:, i - 1] - 2 * (tan[:, i - 1] * v).sum(1, keepdims=True) * v
        v2 = _normalize(tan[:, i] - tL)
        n1 = _normalize(nL - 2 * (nL * v2).sum(1, keepdims=True) * v2)
        keep = vn[:, 0] < 1e-8
        N[:, i] = np.where(keep[:, None], N[:, i - 1], n1)
        Bn[:, i] = np.where(keep[:, None], Bn[:, i - 1], np.cross(tan[:, i], n1))
    off = np.array([-0.5 * width, 0.5 * width])
    return c[:, None] + off[None, :, None, None] * Bn[:, None]


def grid_triangles(rows: int, n: int) -> np.ndarray:
    """Triangles of a rows x n vertex grid: [i0,i1,i2], [i0,i2,i3] per quad."""
    tris = []
    for i in range(rows - 1):
        for j in range(n - 1):
            i0, i1 = i * n + j, (i + 1) * n + j
            i2, i3 = (i + 1) * n + j + 1, i * n + j + 1
            tris.append([i0, i1, i2])
            tris.append([i0, i2, i3])
    return np.asarray(tris, np.int32).reshape(-1, 3)


def edges_from_triangles(tri: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Unique undirected edges (E, 2) [min, max] and their adjacent
    triangles (E, 2) (second = -1 on the boundary), in order of first
    appearance over (t, (i0,i1), (i1,i2), (i2,i0))."""
    tri = np.asarray(tri, np.int64)
    T = tri.shape[0]
    a = tri[:, [0, 1, 2]].reshape(-1)
    b = tri[:, [1, 2, 0]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(tri.max(initial=0)) + 1) + hi
    order = np.lexsort((np.arange(key.size), key))
    ks = key[order]
    start = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    first = order[start]
    size = np.diff(np.r_[start, ks.size])
    second = np.where(size >= 2, order[np.minimum(start + 1, ks.size - 1)], -1)
    o = np.argsort(first)
    first, second = first[o], second[o]
    edges = np.stack([lo[first], hi[first]], 1).astype(np.int32)
    e2t = np.stack([first // 3, np.where(second >= 0, second // 3, -1)], 1).astype(np.int32)
    del T
    return edges, e2t


@dataclass
class Cards:
    verts: np.ndarray      # (V, 3) float32
    tri: np.ndarray        # (T, 3) int32
    tri2cluster: np.ndarray  # (T,) int32
    edges: np.ndarray      # (E, 2) int32
    edge2tri: np.ndarray   # (E, 2) int32
    edge2cluster: np.ndarray  # (E,) int32


def build_cards(strands: np.ndarray, labels: np.ndarray, num_clusters: int, quads: int,
                width: float = 0.01) -> Cards:
    """One quad strip of ``quads`` quads per cluster (clusters without
    strands get no card)."""
    n_v = strands.shape[1]
    sums = np.zeros((num_clusters, n_v, 3))
    np.add.at(sums, labels, strands.astype(np.float64))
    cnt = np.bincount(labels, minlength=num_clusters)
    ids = np.flatnonzero(cnt > 0)
    mean = sums[ids] / cnt[ids, None, None]
    centre = resample_by_arclength(mean, quads + 1)
    rows = strip_rows(centre, width)  # (C, 2, Q+1, 3)
    C = rows.shape[0]
    nv = 2 * (quads + 1)
    verts = rows.reshape(C * nv, 3).astype(np.float32)
    t1 = grid_triangles(2, quads + 1)
    tri = (t1[None] + (np.arange(C) * nv)[:, None, None]).reshape(-1, 3).astype(np.int32)
    t2c = np.repeat(ids.astype(np.int32), t1.shape[0])
    edges, e2t = edges_from_triangles(tri)
    return Cards(verts, tri, t2c, edges, e2t, t2c[e2t[:, 0]].astype(np.int32))


# -------------------------------------------------------------- workloads


@dataclass
class Workload:
    name: str
    H: int
    W: int
    strands: np.ndarray       # (N, n_v, 3)
    labels: np.ndarray        # (N,)
    num_clusters: int
    pos: np.ndarray           # (B, N*n_v, 4) strand vertices, clip space
    edges: np.ndarray         # (N*(n_v-1), 2)
    edge2cluster: np.ndarray  # (N*(n_v-1),)
    cards: Cards
    card_pos: np.ndarray      # (B, Vc, 4)
    mvp: np.ndarray


def strand_edges(num_strands: int, num_vertices: int) -> np.ndarray:
    """Consecutive-vertex edges of every strand (build_clustered_strand_edges,
    3_fitting.py:17-60)."""
    base = np.arange(num_strands * num_vertices, dtype=np.int64).reshape(num_strands, num_vertices)
    return np.stack([base[:, :-1].reshape(-1), base[:, 1:].reshape(-1)], 1).astype(np.int32)


def make_workload(name: str, num_strands: int, num_vertices: int, step: float, seed: int,
                  views: int, res: int, clusters: str, num_clusters: int, quads: int,
                  margin: float = 0.3) -> Workload:
    strands = generate_strands(num_strands, num_vertices, step, seed)
    if clusters == "grid":
        labels = grid_clusters(strands[:, 0])
    else:
        labels = kmeans_clusters(strands[:, 0], num_clusters, seed)
    pts = strands.reshape(-1, 3)
    mvp = ortho_cameras(pts, views, margin)
    pos = project(pts, mvp)
    edges = strand_edges(num_strands, num_vertices)
    e2c = np.repeat(labels, num_vertices - 1).astype(np.int32)
    cards = build_cards(strands, labels, num_clusters, quads)
    return Workload(name, res, res, strands, labels, num_clusters, pos, edges, e2c, cards,
                    project(cards.verts, mvp), mvp)


def small_workload() -> Workload:
    """2,000 strands x 32 vertices (step 0.004, seed 0), 64 grid clusters,
    1 view at 256 x 256, 64 cards x 16 quads."""
    return make_workload("2k strands x 32 verts, 1 view 256^2, K=64, 64 cards x 16 quads",
                         2000, 32, 0.004, 0, 1, 256, "grid", 64, 16)


def full_workload() -> Workload:
    """50,000 strands x 100 vertices (step 0.003, seed 1), 1,024 k-means
    clusters, 4 views at 1024 x 1024, 1,024 cards x 20 quads."""
    return make_workload("50k strands x 100 verts, 4 views 1024^2, K=1024, 1024 cards x 20 quads",
                         50000, 100, 0.003, 1, 4, 1024, "kmeans", 1024, 20)


def adjoint_workload() -> Workload:
    """The strands and cards of ``full_workload`` (50,000 x 100, 1,024
    k-means clusters, 1,024 cards x 20 quads) seen from 8 views at 512 x 512:
    the card-density fitting problem of the hand-written adjoint."""
    return make_workload("50k strands x 100 verts, 8 views 512^2, K=1024, 1024 cards x 20 quads",
                         50000, 100, 0.003, 1, 8, 512, "kmeans", 1024, 20)
