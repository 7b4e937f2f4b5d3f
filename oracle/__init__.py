"""CPU oracle for the pystrns2cards hot path — TEST INFRASTRUCTURE ONLY.

numpy front end over ``liboracle.so`` (built from ``s2c_oracle.c`` by
``oracle/Makefile``), a plain-C, element-for-element restatement of the
reference CUDA kernels (kenji-tojo/strns2cards, pystrns2cards/src/*.cu).
Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
use it, and only as the reference the GPU results are compared against / the
timed CPU path; the product never does.

Functions take and return numpy arrays and follow the reference's *Python*
API semantics (allocation, -1 fill, stable compaction), e.g.
``dda_compute_fragments`` = dda.py:5-76 on the CPU.

Parity pins live in tests/test_oracle_*.py.  Pinned by closed forms / brute
force / invariants: DDA span+fragments, triangle raster, depth test, cluster
masks, bitset RGB/popcount, silhouette marking, compaction, smoothing,
skinning.  The antialias backward is pinned by hand-computable cases (one
straight edge, translation equivariance, rho scaling) — see its tests.
The large-steps section (float64 numpy: Laplacian, system matrix, Cholesky
solve) is pinned by hand-computed matrices and the closed-form spectrum of
a grid's system matrix (tests/test_oracle_largesteps.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_opt_i32p = ctypes.c_void_p  # an int32 array's .ctypes.data, or None
_opt_f32p = ctypes.c_void_p  # a float32 array's .ctypes.data, or None
_I = ctypes.c_int
_F = ctypes.c_float

_SIGS = {
    "orc_dda_compute_span": (_I, [_I, _i32p, _I, _I, _I, _f32p, _I, _i32p, _i32p, _f32p, _f32p]),
    "orc_dda_compute_fragments": (None, [_I, _I, _i32p, _i32p, _f32p, _f32p, _I, _I, _I, _f32p,
                                         _I, _i32p, _i32p, _f32p]),
    "orc_dda_interpolate_attributes": (None, [_I, _I, _I, _I, _f32p, _f32p, _i32p, _f32p]),
    "orc_backward_dda_interpolate_attributes": (None, [_I, _I, _I, _f32p, _I, _f32p, _i32p, _f32p,
                                                       _opt_f32p]),
    "orc_compute_triangle_rects": (_I, [_I, _I, _I, _I, _f32p, _I, _i32p, _i32p, _i32p]),
    "orc_compute_fragments": (None, [_I, _I, _I, _f32p, _I, _i32p, _I, _I, _i32p, _i32p, _i32p,
                                     _f32p]),
    "orc_depth_test": (None, [_I, _I, _I, _I, _i32p, _f32p, _f32p]),
    "orc_filter_valid_fragments": (_I, [_I, _i32p, _f32p, _i32p, _f32p]),
    "orc_interpolate_triangle_attributes": (None, [_I, _I, _I, _I, _f32p, _f32p, _i32p, _f32p]),
    "orc_backward_interpolate_triangle_attributes": (None, [_I, _I, _I, _I, _f32p, _f32p, _i32p,
                                                            _f32p, _opt_f32p]),
    "orc_cluster_mask_from_fragments": (None, [_I, _i32p, _f32p, _i32p, _I, _I, _I, _i32p]),
    "orc_accumulate_bitset_rgb": (None, [_I, _I, _I, _I, _i32p, _f32p, _f32p]),
    "orc_popcount_bitset": (None, [_I, _I, _I, _I, _i32p, _f32p]),
    "orc_mark_discontinuity_edges": (None, [_I, _I, _I, _I, _f32p, _i32p, _I, _i32p, _i32p,
                                            _i32p, _f32p]),
    "orc_compact_valid_ints": (_I, [_I, _i32p, _i32p]),
    "orc_backward_antialiased_cluster_bitset": (None, [_I, _i32p, _f32p, _I, _I, _I, _i32p, _i32p,
                                                       _I, _f32p, _I, _i32p, _f32p, _i32p, _f32p,
                                                       _F, _F, _opt_f32p]),
    "orc_smooth_strands": (None, [_I, _I, _f32p, _f32p, _f32p, _F, _opt_f32p]),
    "orc_skinning": (None, [_I, _I, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _F, _opt_f32p]),
    "orc_dda_cluster_mask": (None, [_I, _I, _I, _I, _f32p, _I, _i32p, _i32p, _I, _i32p, _I]),
    "orc_triangle_cluster_mask": (None, [_I, _I, _I, _I, _f32p, _I, _i32p, _i32p, _I, _i32p]),
    "orc_card_coverage": (None, [_I, _I, _I, _I, _f64p, _I, _i32p, _I, _opt_i32p, _opt_i32p,
                                 _opt_i32p, _I, _I, _f64p]),
    "orc_card_coverage_grad": (None, [_I, _I, _I, _I, _f64p, _I, _i32p, _f64p, _f64p]),
    "orc_bishop_frame": (None, [_I, _I, _f32p, _f32p, _f32p, _f32p]),
    "orc_resample_strands_by_arclength": (None, [_I, _i32p, _f32p, _I, _f32p]),
}


def build(force: bool = False) -> str:
    """Compiles liboracle.so with oracle/Makefile (gcc)."""
    src = os.path.join(HERE, "s2c_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ DDA

def dda_compute_span(resolution, pos, edges, edge_ids=None):
    """dda.cu:89-122 → (prefix, slopes, spans, total).  Entries of slopes /
    spans for zero-length edges are NaN (the reference leaves them unset)."""
    H, W = resolution
    pos, edges = _f(pos), _i(edges)
    B, V, _ = pos.shape
    E = edges.shape[0]
    ids = _i(np.arange(B * E) if edge_ids is None else edge_ids)
    P = ids.shape[0]
    prefix = np.zeros(P, np.int32)
    slopes = np.full((P, 2), np.nan, np.float32)
    spans = np.full((P, 4), np.nan, np.float32)
    total = lib().orc_dda_compute_span(P, ids, H, W, V, pos, E, edges, prefix, slopes, spans)
    return prefix, slopes, spans, int(total)


def dda_compute_fragments(resolution, pos, edges, edge_ids=None):
    """dda.py:5-76: (frag_pix (N,3) int32, frag_attrs (N,4) float32), stable."""
    H, W = resolution
    pos, edges = _f(pos), _i(edges)
    B, V, _ = pos.shape
    E = edges.shape[0]
    ids = _i(np.arange(B * E) if edge_ids is None else edge_ids)
    prefix, slopes, spans, n = dda_compute_span(resolution, pos, edges, ids)
    fp = np.full((n, 3), -1, np.int32)
    fa = np.zeros((n, 4), np.float32)
    lib().orc_dda_compute_fragments(ids.shape[0], n, prefix, ids, slopes, spans, H, W, V, pos, E,
                                    edges, fp, fa)
    return filter_valid_fragments(fp, fa)


def dda_interpolate_attributes(rast, attr, edges):
    rast, attr, edges = _f(rast), _f(attr), _i(edges)
    B, H, W, _ = rast.shape
    C = attr.shape[1]
    img = np.zeros((B, H, W, C), np.float32)
    lib().orc_dda_interpolate_attributes(B, H, W, C, rast, attr, edges, img)
    return img


def _abs_out(like, with_abs):
    """(array, pointer) of a zero-filled Σ|term| output, or (None, None)."""
    if not with_abs:
        return None, None
    a = np.zeros_like(like)
    return a, a.ctypes.data


def backward_dda_interpolate_attributes(rast, d_image, edges, num_vertices, with_abs=False):
    """d_attr; with ``with_abs`` also each element's sum of |term|."""
    rast, d_image, edges = _f(rast), _f(d_image), _i(edges)
    B, H, W, _ = rast.shape
    C = d_image.shape[3]
    d_attr = np.zeros((num_vertices, C), np.float32)
    ab, abp = _abs_out(d_attr, with_abs)
    lib().orc_backward_dda_interpolate_attributes(B, H, W, rast, C, d_image, edges, d_attr, abp)
    return (d_attr, ab) if with_abs else d_attr


def dda_cluster_mask(resolution, pos, edges, edge2cluster, cluster_count, threads=1):
    """dda.py:5-76 + rasterize.py:170-214 as one pass; B = pos.shape[0]."""
    H, W = resolution
    pos, edges, e2c = _f(pos), _i(edges), _i(edge2cluster)
    B, V, _ = pos.shape
    S = (cluster_count + 31) // 32
    bits = np.zeros((B, H, W, S), np.int32)
    lib().orc_dda_cluster_mask(B, H, W, V, pos, edges.shape[0], edges, e2c, S, bits, int(threads))
    return bits


# ------------------------------------------------------------- triangles

def compute_triangle_rects(resolution, pos, tri):
    H, W = resolution
    pos, tri = _f(pos), _i(tri)
    B, V, _ = pos.shape
    T = tri.shape[0]
    rects = np.zeros((B * T, 4), np.int32)
    prefix = np.zeros(B * T, np.int32)
    n = lib().orc_compute_triangle_rects(H, W, B, V, pos, T, tri, rects, prefix)
    return rects, prefix, int(n)


def compute_fragments(resolution, pos, tri):
    """rasterize.py:6-53 on the CPU, stable order."""
    H, W = resolution
    pos, tri = _f(pos), _i(tri)
    B, V, _ = pos.shape
    T = tri.shape[0]
    rects, prefix, n = compute_triangle_rects(resolution, pos, tri)
    fp = np.full((n, 3), -1, np.int32)
    fa = np.zeros((n, 4), np.float32)
    lib().orc_compute_fragments(H, W, V, pos, T, tri, B * T, n, prefix, rects, fp, fa)
    return filter_valid_fragments(fp, fa)


def filter_valid_fragments(fp, fa):
    fp, fa = _i(fp), _f(fa)
    n = fp.shape[0]
    fp_out = np.zeros((n, 3), np.int32)
    fa_out = np.zeros((n, 4), np.float32)
    m = lib().orc_filter_valid_fragments(n, fp, fa, fp_out, fa_out)
    return fp_out[:m].copy(), fa_out[:m].copy()


def depth_test(B, resolution, frag_pix, frag_attrs):
    H, W = resolution
    fp, fa = _i(frag_pix), _f(frag_attrs)
    rast = np.zeros((B, H, W, 4), np.float32)
    lib().orc_depth_test(B, H, W, fp.shape[0], fp, fa, rast)
    return rast


def rasterize(resolution, pos, tri):
    fp, fa = compute_fragments(resolution, pos, tri)
    return depth_test(np.asarray(pos).shape[0], resolution, fp, fa)


def interpolate_triangle_attributes(rast, attr, tri):
    rast, attr, tri = _f(rast), _f(attr), _i(tri)
    B, H, W, _ = rast.shape
    C = attr.shape[1]
    img = np.zeros((B, H, W, C), np.float32)
    lib().orc_interpolate_triangle_attributes(B, H, W, C, rast, attr, tri, img)
    return img


def backward_interpolate_triangle_attributes(rast, d_image, tri, num_vertices, with_abs=False):
    """d_attr; with ``with_abs`` also each element's sum of |term|."""
    rast, d_image, tri = _f(rast), _f(d_image), _i(tri)
    B, H, W, _ = rast.shape
    C = d_image.shape[3]
    d_attr = np.zeros((num_vertices, C), np.float32)
    ab, abp = _abs_out(d_attr, with_abs)
    lib().orc_backward_interpolate_triangle_attributes(B, H, W, C, rast, d_image, tri, d_attr, abp)
    return (d_attr, ab) if with_abs else d_attr


def cluster_mask_from_fragments(resolution, frag_pix, frag_attrs, prim2cluster, cluster_count=None,
                                B=None):
    """rasterize.py:170-214; B defaults to frag_pix[:,0].max()+1."""
    H, W = resolution
    fp, fa, p2c = _i(frag_pix), _f(frag_attrs), _i(prim2cluster)
    if B is None:
        B = int(fp[:, 0].max()) + 1
    if cluster_count is None:
        cluster_count = int(p2c.max()) + 1
    S = (cluster_count + 31) // 32
    bits = np.zeros((B, H, W, S), np.int32)
    lib().orc_cluster_mask_from_fragments(fp.shape[0], fp, fa, p2c, H, W, S, bits)
    return bits


def triangle_cluster_mask(resolution, pos, tri, tri2cluster, cluster_count):
    H, W = resolution
    pos, tri, t2c = _f(pos), _i(tri), _i(tri2cluster)
    B, V, _ = pos.shape
    S = (cluster_count + 31) // 32
    bits = np.zeros((B, H, W, S), np.int32)
    lib().orc_triangle_cluster_mask(B, H, W, V, pos, tri.shape[0], tri, t2c, S, bits)
    return bits


# ---------------------------------------------------------------- bitset

def accumulate_bitset_rgb(bitset, rgb):
    bits, rgb = _i(bitset), _f(rgb)
    B, H, W, S = bits.shape
    out = np.zeros((B, H, W, 3), np.float32)
    lib().orc_accumulate_bitset_rgb(B, H, W, S, bits, rgb, out)
    return out


def popcount_bitset(bitset):
    bits = _i(bitset)
    B, H, W, S = bits.shape
    out = np.zeros((B, H, W), np.float32)
    lib().orc_popcount_bitset(B, H, W, S, bits, out)
    return out


# ------------------------------------------------------------- antialias

def mark_discontinuity_edges(resolution, pos, tri, edges, edge2tri):
    """antialias.py:8-52 → (prim_ids, normals)."""
    H, W = resolution
    pos, tri, edges, e2t = _f(pos), _i(tri), _i(edges), _i(edge2tri)
    B, V, _ = pos.shape
    E = edges.shape[0]
    prim = np.full(B * E, -1, np.int32)
    normals = np.zeros((B * E, 2), np.float32)
    lib().orc_mark_discontinuity_edges(H, W, B, V, pos, tri, E, edges, e2t, prim, normals)
    return compact_valid_ints(prim), normals


def compact_valid_ints(a):
    a = _i(a)
    out = np.zeros_like(a)
    n = lib().orc_compact_valid_ints(a.shape[0], a, out)
    return out[:n].copy()


def backward_antialiased_cluster_bitset(frag_pix, frag_attrs, bitset, target, pos, edges, normals,
                                        edge2cluster, kernel_radius=7.0, rho=1.0, with_abs=False):
    """d_pos; with ``with_abs`` also each element's sum of |per-sample term|."""
    fp, fa = _i(frag_pix), _f(frag_attrs)
    bits, tgt = _i(bitset), _i(target)
    pos, edges, normals, e2c = _f(pos), _i(edges), _f(normals), _i(edge2cluster)
    B, H, W, S = bits.shape
    V = pos.shape[1]
    d_pos = np.zeros_like(pos)
    ab, abp = _abs_out(d_pos, with_abs)
    lib().orc_backward_antialiased_cluster_bitset(fp.shape[0], fp, fa, H, W, S, bits, tgt, V, pos,
                                                  edges.shape[0], edges, normals, e2c, d_pos,
                                                  float(kernel_radius), float(rho), abp)
    return (d_pos, ab) if with_abs else d_pos


def antialiased_cluster_bitset_grad(bitset, target, pos, tri, edges, edge2tri, edge2cluster,
                                    kernel_radius=7.0, rho=1.0, with_abs=False):
    """d_pos of antialias.py:66-130 (the gradient its loss carries); with
    ``with_abs`` also each element's sum of |per-sample term|."""
    H, W = np.asarray(target).shape[1:3]
    prim, normals = mark_discontinuity_edges((H, W), pos, tri, edges, edge2tri)
    fp, fa = dda_compute_fragments((H, W), pos, edges, prim)
    return backward_antialiased_cluster_bitset(fp, fa, bitset, target, pos, edges, normals,
                                               edge2cluster, kernel_radius, rho, with_abs)


# ------------------------------------------------------------------ curves

def compute_arclength(strands):
    """curve_utils.py:167-180 in float32."""
    x = _f(strands)
    seg = np.sqrt(((x[:, 1:] - x[:, :-1]) ** 2).sum(-1, dtype=np.float32)).astype(np.float32)
    return np.concatenate([np.zeros((x.shape[0], 1), np.float32), np.cumsum(seg, 1, dtype=np.float32)],
                          1).astype(np.float32)


def smooth_strands(strands, arclength, beta, with_abs=False):
    """Smoothed strands; with ``with_abs`` also Σ w|x| / Σ w per element."""
    x, s = _f(strands), _f(arclength)
    N, P, _ = x.shape
    out = np.zeros_like(x)
    ab, abp = _abs_out(out, with_abs)
    lib().orc_smooth_strands(N, P, x, s, out, float(beta), abp)
    return (out, ab) if with_abs else out


def bishop_frame(strands):
    """(T, N, B) rotation-minimizing frames, each (N, S, 3) float32."""
    x = _f(strands)
    N, S, _ = x.shape
    T = np.zeros_like(x)
    Nr = np.zeros_like(x)
    Bn = np.zeros_like(x)
    lib().orc_bishop_frame(N, S, x, T, Nr, Bn)
    return T, Nr, Bn


def skinning(strands, arclength, handle, canonical, R, beta, with_abs=False):
    """Skinned strands; with ``with_abs`` also the weighted sum of |term|
    per element."""
    x, s, h, c, r = _f(strands), _f(arclength), _f(handle), _f(canonical), _f(R)
    N, P, _ = x.shape
    out = np.zeros_like(x)
    ab, abp = _abs_out(out, with_abs)
    lib().orc_skinning(N, P, x, s, h, c, r, out, float(beta), abp)
    return (out, ab) if with_abs else out


# ------------------------------------------------------ card coverage (adjoint)


def _d(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def card_coverage(resolution, pos, tri, sel=None, window=None) -> np.ndarray:
    """Area-sampled coverage (B, H, W) float64 of the triangles (all, or the
    indices ``sel``) of each view; ``window`` = (y0[B], x0[B], h, w) restricts
    the output to an h x w window per view."""
    H, W = resolution
    pos = _d(pos)
    tri = _i(tri)
    B, V = pos.shape[:2]
    if window is None:
        y0 = x0 = None
        h, w = H, W
    else:
        y0, x0, h, w = window
        y0, x0 = _i(y0), _i(x0)
    out = np.zeros((B, h, w), np.float64)
    sel_a = None if sel is None else _i(sel)
    lib().orc_card_coverage(B, H, W, V, pos, tri.shape[0],
                            tri, 0 if sel_a is None else sel_a.size,
                            None if sel_a is None else sel_a.ctypes.data,
                            None if y0 is None else y0.ctypes.data,
                            None if x0 is None else x0.ctypes.data, h, w, out)
    return out


def card_coverage_grad(resolution, pos, tri, d_cov) -> np.ndarray:
    """d loss / d pos (B, V, 4) float64 given d loss / d coverage."""
    H, W = resolution
    pos = _d(pos)
    tri = _i(tri)
    B, V = pos.shape[:2]
    out = np.zeros((B, V, 4), np.float64)
    lib().orc_card_coverage_grad(B, H, W, V, pos, tri.shape[0], tri, _d(d_cov), out)
    return out


def _screen(pos, H, W):
    x = 0.5 * (1.0 + pos[..., 0] / pos[..., 3]) * W
    y = 0.5 * (1.0 - pos[..., 1] / pos[..., 3]) * H
    return x, y


def coverage_loss_difference(resolution, pos, tri, target, cov, delta) -> float:
    """(L(pos + delta) - L(pos - delta)) / 2 for L = sum (C - target)^2, in
    float64, where ``cov`` = card_coverage(pos) and ``delta`` (B, V, 4) is
    non-zero on a few vertices: only the triangles touching them are
    rendered again, inside a per-view window around them."""
    H, W = resolution
    pos = _d(pos)
    delta = _d(delta)
    tri = _i(tri)
    verts = np.flatnonzero(np.any(delta != 0, axis=(0, 2)))
    sel = np.flatnonzero(np.isin(tri, verts).any(1)).astype(np.int32)
    if sel.size == 0:
        return 0.0
    B = pos.shape[0]
    vs = np.unique(tri[sel])
    xs, ys = [], []
    for p in (pos - delta, pos, pos + delta):
        x, y = _screen(p[:, vs], H, W)
        xs.append(x)
        ys.append(y)
    xs, ys = np.concatenate(xs, 1), np.concatenate(ys, 1)
    x0 = np.clip(np.floor(xs.min(1)) - 1, 0, W - 1).astype(np.int32)
    y0 = np.clip(np.floor(ys.min(1)) - 1, 0, H - 1).astype(np.int32)
    x1 = np.clip(np.floor(xs.max(1)) + 2, 1, W).astype(np.int32)
    y1 = np.clip(np.floor(ys.max(1)) + 2, 1, H).astype(np.int32)
    h, w = int((y1 - y0).max()), int((x1 - x0).max())
    win = (y0, x0, h, w)
    old = card_coverage(resolution, pos, tri, sel, win)
    e = np.zeros((B, h, w))
    for b in range(B):
        hb = min(h, H - y0[b])
        wb = min(w, W - x0[b])
        e[b, :hb, :wb] = (np.asarray(cov, np.float64)[b, y0[b]:y0[b] + hb, x0[b]:x0[b] + wb] -
                          np.asarray(target, np.float64)[b, y0[b]:y0[b] + hb, x0[b]:x0[b] + wb])
    lp = ((e + card_coverage(resolution, pos + delta, tri, sel, win) - old) ** 2).sum()
    lm = ((e + card_coverage(resolution, pos - delta, tri, sel, win) - old) ** 2).sum()
    return 0.5 * (lp - lm)


def resample_strands_by_arclength(num_samples, points, target):
    """(N, target, 3) float32 arclength resampling (hair_loader.cpp:261-307)."""
    ns = _i(np.asarray(num_samples).ravel())
    out = np.zeros((ns.size, int(target), 3), np.float32)
    lib().orc_resample_strands_by_arclength(ns.size, ns, _f(points), int(target), out)
    return out


# ------------------------------------------------ large steps (grid meshes)
# The reference's third-party solve: largesteps 0.2.2 (parameterize.py's
# to_differential(L, v) = L @ v and from_differential(L, u), a cholespy
# 2.1.0 sparse Cholesky solve of L x = u; requirements-detail.txt:1,14),
# neither installed here.  Restated in float64 with numpy/scipy: the
# published algorithm is a Cholesky factorization and two triangular solves.

def wisp_grid_edges(S: int, N: int):
    """h_edges and w_edges of one (S, N) wisp, loop for loop as
    texture_utils.py:566-584 (Python loops: small cases only)."""
    h = [(i * N + j, i * N + j + 1) for i in range(S) for j in range(N - 1)]
    w = [(i * N + j, (i + 1) * N + j) for i in range(S - 1) for j in range(N)]
    return np.array(h, np.int64).reshape(-1, 2), np.array(w, np.int64).reshape(-1, 2)


def combined_grid_edges(shapes):
    """Edges of several (S, N) wisps concatenated with running vertex
    offsets, as combine_wisps_with_cluster_labels (texture_utils.py:634-655).
    Returns (num_verts, h_edges, w_edges)."""
    hs, ws, off = [], [], 0
    for S, N in shapes:
        h, w = wisp_grid_edges(S, N)
        hs.append(h + off)
        ws.append(w + off)
        off += S * N
    cat = lambda a: np.concatenate(a) if a else np.zeros((0, 2), np.int64)  # noqa: E731
    return off, cat(hs), cat(ws)


def laplacian_edges(V: int, edges) -> np.ndarray:
    """Dense float64 uniform Laplacian, as optimizer.py:70-99: the distinct
    directed pairs of the symmetrized edge list get -1, each pair adds +1 at
    its first vertex's diagonal (duplicates summed, as coalesce does)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    adj = np.unique(np.concatenate([e, e[:, ::-1]]), axis=0)
    L = np.zeros((V, V))
    np.add.at(L, (adj[:, 0], adj[:, 1]), -1.0)
    np.add.at(L, (adj[:, 0], adj[:, 0]), 1.0)
    return L


def largesteps_matrix(V: int, h_edges, w_edges, lam: float, w_weight: float) -> np.ndarray:
    """M = I + lam * (L_h + w_weight * L_w) (3_fitting.py:347-356)."""
    return np.eye(V) + lam * (laplacian_edges(V, h_edges) + w_weight * laplacian_edges(V, w_edges))


def to_differential(M: np.ndarray, x) -> np.ndarray:
    """u = M x (largesteps to_differential)."""
    return M @ np.asarray(x, np.float64)


def from_differential(M: np.ndarray, u) -> np.ndarray:
    """x = M^{-1} u through the Cholesky factor M = C C^T (largesteps
    from_differential with its default Cholesky method)."""
    from scipy.linalg import solve_triangular

    C = np.linalg.cholesky(M)
    y = solve_triangular(C, np.asarray(u, np.float64), lower=True)
    return solve_triangular(C.T, y, lower=False)


# --------------------------------------------------------------- card meshes

def strand_to_strip_mesh(strand, res_w=2, width=0.5, tangent_rotations_deg=(0.0,)):
    """Strips of one strand, as 2_texture.py:17-96: central-difference
    tangents, a root frame from the change of tangent (or a fixed axis when
    it is below 1e-4), rotated about the root tangent by each angle (Rodrigues),
    propagated by the double-reflection rotation-minimizing frame, and the
    strip swept along the binormal at res_w offsets across `width`.  float32
    like the reference's arrays (its angle terms are float64 and rounded
    into the float32 frame).  Returns a list of (res_w, S, 3) float32."""
    p = np.asarray(strand, np.float32)
    S = p.shape[0]
    d = p[1:] - p[:-1]
    t = np.zeros_like(p)
    t[:-1] += d
    t[1:] += d
    t /= np.maximum(np.linalg.norm(t, axis=1, keepdims=True), np.float32(1e-8))
    T0 = t[0]
    dT = t[1] - t[0]
    if np.linalg.norm(dT) < 1e-4:
        n = np.array([0.0, 1.0, 0.0], np.float32)
        if abs(np.dot(T0, n)) > 0.9:
            n = np.array([0.0, 0.0, 1.0], np.float32)
    else:
        n = -dT / np.linalg.norm(dT)
    b = np.cross(T0, n)
    b /= max(np.linalg.norm(b), np.float32(1e-8))
    n = np.cross(b, T0)
    off = np.linspace(-0.5 * width, 0.5 * width, res_w, dtype=np.float32)
    strips = []
    for deg in tangent_rotations_deg:
        th = np.deg2rad(deg)
        N0 = np.cos(th) * n + np.sin(th) * np.cross(T0, n) + (1.0 - np.cos(th)) * np.dot(T0, n) * T0
        Ni = N0.astype(np.float32)
        Bs = np.zeros_like(p)
        Bs[0] = np.cross(T0, N0)
        for i in range(1, S):
            v = p[i] - p[i - 1]
            vn = np.linalg.norm(v)
            if vn < 1e-8:
                Bs[i] = Bs[i - 1]
                continue
            v = v / vn
            nl = Ni - 2 * np.dot(Ni, v) * v
            tl = t[i - 1] - 2 * np.dot(t[i - 1], v) * v
            v2 = t[i] - tl
            v2 /= max(np.linalg.norm(v2), np.float32(1e-8))
            n1 = nl - 2 * np.dot(nl, v2) * v2
            n1 /= max(np.linalg.norm(n1), np.float32(1e-8))
            Ni = n1
            Bs[i] = np.cross(t[i], n1)
        strips.append((p[None] + off[:, None, None] * Bs[None]).astype(np.float32))
    return strips


def grid_mesh(S: int, N: int, uv_rect=None, margin_top=1e-4):
    """Faces, quads, h_edges, w_edges (and uvs) of one (S, N) wisp, as
    texture_utils.py:525-599 (Python loops: small cases)."""
    faces, quads = [], []
    for i in range(S - 1):
        for j in range(N - 1):
            a, b, c, d = i * N + j, (i + 1) * N + j, (i + 1) * N + j + 1, i * N + j + 1
            faces += [[a, b, c], [a, c, d]]
            quads.append([a, b, c, d])
    h, w = wisp_grid_edges(S, N)
    out = dict(faces=np.array(faces, np.int32).reshape(-1, 3),
               quads=np.array(quads, np.int32).reshape(-1, 4),
               h_edges=h.astype(np.int32), w_edges=w.astype(np.int32))
    if uv_rect is not None:
        u0, v0, u1, v1 = uv_rect
        uu, vv = np.meshgrid(np.linspace(u0, u1, S), np.linspace(v0, v1 - margin_top, N),
                             indexing="ij")
        out["uvs"] = np.stack([uu, vv], -1).reshape(-1, 2)
    return out


def combine_grid_meshes(shapes, rects, grid2cluster):
    """texture_utils.py:597-655: the wisps' meshes concatenated with running
    vertex offsets and a cluster label per face.  Returns (faces, uvs,
    face2cluster, quads, h_edges, w_edges)."""
    parts = {k: [] for k in ("faces", "uvs", "f2c", "quads", "h_edges", "w_edges")}
    off = 0
    for (S, N), rect, c in zip(shapes, rects, grid2cluster):
        m = grid_mesh(S, N, rect)
        for k in ("faces", "quads", "h_edges", "w_edges"):
            parts[k].append(m[k] + off)
        parts["uvs"].append(m["uvs"])
        parts["f2c"].append(np.full(m["faces"].shape[0], c))
        off += S * N
    cat = np.concatenate
    return (cat(parts["faces"]), cat(parts["uvs"]), cat(parts["f2c"]), cat(parts["quads"]),
            cat(parts["h_edges"]), cat(parts["w_edges"]))


def clustered_strand_edges(counts, cluster_ids):
    """Edges between consecutive points of each strand and the strand's
    cluster per edge, as 3_fitting.py:17-60 (points concatenated)."""
    edges, e2c, off = [], [], 0
    for n, c in zip(counts, cluster_ids):
        edges += [[off + i, off + i + 1] for i in range(n - 1)]
        e2c += [c] * (n - 1)
        off += n
    return (np.array(edges, np.int32).reshape(-1, 2), np.array(e2c, np.int32),
            max([c + 1 for c in cluster_ids], default=0))


# ------------------------------------------------------- MSAA texture patch

def blur3x3_reflect(a, k):
    """Separable 3-tap blur, reflect borders, along x then y, each output
    (k0 a[-1] + k1 a[0]) + k2 a[+1] in float32 (kornia's gaussian_blur2d
    with a 3 x 3 kernel, texture_utils.py:250-254, without its conv's
    unspecified summation order)."""
    a = np.asarray(a, np.float32)
    k0, k1, k2 = (np.float32(v) for v in k)
    H, W = a.shape
    xl = np.abs(np.arange(W) - 1)
    xr = np.arange(W) + 1
    xr[xr >= W] = 2 * W - 2 - xr[xr >= W]
    t = (k0 * a[:, xl] + k1 * a) + k2 * a[:, xr]
    yl = np.abs(np.arange(H) - 1)
    yr = np.arange(H) + 1
    yr[yr >= H] = 2 * H - 2 - yr[yr >= H]
    return ((k0 * t[yl, :] + k1 * t) + k2 * t[yr, :]).astype(np.float32)


def msaa_patch(pos, tri, attr, H, W, dx, dy, weights, kernel3):
    """One wisp patch of texture_utils.py:203-254: per jitter s a separate
    rasterize + interpolate of the (V, 3) NDC mesh moved by (dx[s], dy[s])
    (float32), rgb averaged over covered samples ((0, 1, 0) where none),
    mapped by 0.5 (rgb + 1), alpha the weighted coverage, blurred.
    Returns (rgb (H, W, 3), alpha (H, W)) float32, before the flip."""
    pos = np.asarray(pos, np.float32)
    tri = np.asarray(tri, np.int32)
    attr = np.asarray(attr, np.float32)
    V = pos.shape[0]
    acc = np.zeros((H, W, 3), np.float32)
    cnt = np.zeros((H, W), np.float32)
    al = np.zeros((H, W), np.float32)
    for s in range(len(dx)):
        p = np.zeros((1, V, 4), np.float32)
        p[0, :, 0] = pos[:, 0] + np.float32(dx[s])
        p[0, :, 1] = pos[:, 1] + np.float32(dy[s])
        p[0, :, 2] = pos[:, 2]
        p[0, :, 3] = 1.0
        rast = rasterize((H, W), p, tri)
        img = interpolate_triangle_attributes(rast, attr, tri)[0]
        v = (rast[0, ..., 3] > 0).astype(np.float32)
        acc += img * v[..., None]
        cnt += v
        al += np.float32(weights[s]) * v
    rgb = acc / np.maximum(cnt, np.float32(1))[..., None]
    rgb[cnt == 0] = np.array([0, 1, 0], np.float32)
    rgb = np.float32(0.5) * (rgb + np.float32(1))
    return rgb.astype(np.float32), blur3x3_reflect(al, kernel3)


def patch_mask(pts, rows, cols, H, W):
    """One wisp's mask of texture_utils.py:346-358: dda_compute_fragments
    + depth_test + dda_interpolate_attributes of a constant 1 over the
    edges of build_edge_indices_from_wisp, flipped, clamped, bytes.
    pts (rows*cols, 3) float32 NDC points."""
    pts = np.asarray(pts, np.float32)
    pos = np.ones((1, pts.shape[0], 4), np.float32)
    pos[0, :, :3] = pts
    idx = np.arange(rows * cols).reshape(rows, cols)
    edges = np.stack([idx[:, :-1].reshape(-1), idx[:, 1:].reshape(-1)], 1).astype(np.int32)
    fp, fa = dda_compute_fragments((H, W), pos, edges)
    rast = depth_test(1, (H, W), fp, fa)
    m = dda_interpolate_attributes(rast, np.ones((pts.shape[0], 1), np.float32), edges)[0, ..., 0]
    m = np.clip(m[::-1], 0, 1)
    return (m * 255).astype(np.uint8)



def from_differential_sparse(V: int, h_edges, w_edges, lam: float, w_weight: float, u) -> np.ndarray:
    """x = M^{-1} u in float64 through scipy's sparse LU of the same M
    (laplacian_edges restated on sparse matrices): the check at sizes the
    dense Cholesky cannot hold."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    def lap(e):
        e = np.asarray(e, np.int64).reshape(-1, 2)
        adj = np.unique(np.concatenate([e, e[:, ::-1]]), axis=0)
        A = sp.coo_matrix((np.ones(len(adj)), (adj[:, 0], adj[:, 1])), shape=(V, V)).tocsr()
        return sp.diags(np.asarray(A.sum(1)).ravel()) - A

    M = (sp.eye(V) + lam * (lap(h_edges) + w_weight * lap(w_edges))).tocsc()
    return spla.splu(M).solve(np.asarray(u, np.float64))
