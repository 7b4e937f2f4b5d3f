"""Strand files, card topology and mesh output: the host-side helpers of the
reference package (pystrns2cards/__init__.py:19-247, src/hair_loader.cpp,
src/mesh_utils.cpp).

* ``load_strands`` reads the reference's three binary formats by extension
  (pystrns2cards.cpp:788-816): ``.bin`` (count, then per strand a count and
  7 floats per point: position, normal, label; strands with a segment
  shorter than 1e-4 are dropped), ``.data`` (USC: count, then per strand a
  count and its points; one-point strands skipped) and ``.hair`` (either the
  "HAIR" header variant or the headerless ZJU layout of counts and points),
  keeping strands of at least 2 points (hair_loader.cpp:15-256).
* ``resample_strands_by_arclength`` runs on the GPU (two HIP kernels with
  the reference's fp32 arithmetic and upper-bound search).
* ``build_edges_from_triangles`` returns the undirected edges of a triangle
  list in first-appearance order (the reference's order comes from a hash
  map and is implementation-defined, mesh_utils.cpp:37-79) with the first
  two adjacent triangles of each edge (-1 when there is one).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _core

__all__ = ["build_edges_from_triangles", "load_strands", "resample_strands_by_arclength",
           "load_strands_arc_length", "save_obj", "load_npz_as_strands",
           "load_npz_as_strands_arc_length", "build_edge_indices_from_wisp",
           "resample_single_strand_by_arclength"]


# ------------------------------------------------------------------ topology


def build_edges_from_triangles(tri: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(edges (E, 2) int32 with v0 < v1, edge2tri (E, 2) int32) of a (T, 3)
    int32 triangle list (reference: __init__.py:19-42)."""
    if not isinstance(tri, np.ndarray):
        raise TypeError("Input must be a NumPy ndarray.")
    if tri.dtype != np.int32:
        raise TypeError("Triangle indices must be int32.")
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("Input triangle array must have shape (T, 3).")
    T = tri.shape[0]
    if T == 0:
        return np.zeros((0, 2), np.int32), np.zeros((0, 2), np.int32)
    # the three edges of triangle t, in the order (0,1), (1,2), (2,0)
    e = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], 1).reshape(-1, 2)
    e = np.sort(e, axis=1).astype(np.int64)
    owner = np.repeat(np.arange(T, dtype=np.int64), 3)
    key = e[:, 0] * (int(e.max()) + 1) + e[:, 1]
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")          # first-appearance order
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    eid = rank[inv]                                   # edge id of every triangle side
    edges = e[first[order]].astype(np.int32)
    # first two triangles per edge, by triangle order
    srt = np.lexsort((owner, eid))
    eid_s, own_s = eid[srt], owner[srt]
    start = np.searchsorted(eid_s, np.arange(order.size))
    cnt = np.bincount(eid_s, minlength=order.size)
    e2t = np.full((order.size, 2), -1, np.int32)
    e2t[:, 0] = own_s[start]
    two = cnt >= 2
    e2t[two, 1] = own_s[start[two] + 1]
    return edges, e2t


def build_edge_indices_from_wisp(wisp) -> np.ndarray:
    """(N*(S-1), 2) edges joining consecutive points of each strand of an
    (N, S, 3) wisp (reference: __init__.py:165-175)."""
    N, S = np.asarray(wisp).shape[:2]
    idx = np.arange(N * S).reshape(N, S)
    return np.stack([idx[:, :-1].reshape(-1), idx[:, 1:].reshape(-1)], axis=1)


# ------------------------------------------------------------------ loaders


def _strands_bin(buf: bytes) -> list[np.ndarray]:
    out, pos = [], 0
    n = int(np.frombuffer(buf, np.int32, 1, pos)[0])
    pos += 4
    for _ in range(n):
        m = int(np.frombuffer(buf, np.int32, 1, pos)[0])
        pos += 4
        rec = np.frombuffer(buf, np.float32, 7 * m, pos).reshape(m, 7)
        pos += 28 * m
        p = rec[:, :3].copy()
        if m > 1 and (np.linalg.norm(p[:-1] - p[1:], axis=1) < np.float32(1e-4)).any():
            continue
        out.append(p)
    return out


def _strands_usc(buf: bytes) -> list[np.ndarray]:
    out, pos = [], 0
    n = int(np.frombuffer(buf, np.int32, 1, pos)[0])
    pos += 4
    for _ in range(n):
        m = int(np.frombuffer(buf, np.int32, 1, pos)[0])
        pos += 4
        if m == 1:
            pos += 12
            continue
        out.append(np.frombuffer(buf, np.float32, 3 * m, pos).reshape(m, 3).copy())
        pos += 12 * m
    return out


def _strands_hair(buf: bytes) -> list[np.ndarray]:
    if buf[:4] == b"HAIR":
        n, count, bits = (int(v) for v in np.frombuffer(buf, np.uint32, 3, 4))
        pos = 16
        if not bits & 2:
            raise ValueError("load_strands: HAIR file missing points section.")
        default_seg = 0
        if not bits & 1:
            default_seg = int(np.frombuffer(buf, np.uint32, 1, pos)[0])
            pos += 4
        # the reference skips 128 - 3 (or - 4) more 32-bit words of header
        pos += 4 * (128 - (3 if default_seg else 4))
        if default_seg == 0:
            seg = np.frombuffer(buf, np.uint16, n, pos).astype(np.int64) + 1
            pos += 2 * n
        else:
            seg = np.full(n, (default_seg & 0xFFFF) + 1, np.int64)
    else:
        n, count = (int(v) for v in np.frombuffer(buf, np.uint32, 2, 0))
        seg = np.frombuffer(buf, np.uint16, n, 8).astype(np.int64)
        pos = 8 + 2 * n
    pts = np.frombuffer(buf, np.float32, 3 * count, pos).reshape(count, 3)
    ends = np.cumsum(seg)
    return [pts[e - s:e].copy() for s, e in zip(seg, ends)]


def load_strands(filepath: str, return_raw: bool = False):
    """(N,) object array of (M_i, 3) float32 strands from a ``.bin``,
    ``.data`` or ``.hair`` file (reference: __init__.py:44-73); with
    ``return_raw`` also (num_samples (N,), points (sum M_i, 3))."""
    if filepath.endswith(".bin"):
        parse = _strands_bin
    elif filepath.endswith(".data"):
        parse = _strands_usc
    elif filepath.endswith(".hair"):
        parse = _strands_hair
    else:
        raise ValueError("load_strands: unknown file extension.")
    with open(filepath, "rb") as f:
        strands = [s for s in parse(f.read()) if s.shape[0] >= 2]
    num_samples = np.array([s.shape[0] for s in strands], np.int32)
    points = np.concatenate(strands, 0) if strands else np.zeros((0, 3), np.float32)
    arr = np.empty(len(strands), dtype=object)
    for i, s in enumerate(strands):
        arr[i] = s
    if return_raw:
        return arr, num_samples, points
    return arr


def resample_strands_by_arclength(num_samples: np.ndarray, points: np.ndarray,
                                  target_num_samples: int = 100) -> np.ndarray:
    """(N, target, 3) float32: every strand resampled to target points evenly
    spaced in arclength, the last point kept exactly (reference:
    __init__.py:75-95, hair_loader.cpp:261-307), computed on the GPU."""
    ns = np.ascontiguousarray(np.asarray(num_samples).ravel(), dtype=np.int32)
    pts = np.ascontiguousarray(np.asarray(points), dtype=np.float32)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must have shape (P, 3)")
    if int(target_num_samples) < 2:
        raise ValueError("target_num_samples must be >= 2")
    if ns.size and int(ns.min()) < 2:
        raise ValueError("every strand needs at least 2 points")
    if int(ns.astype(np.int64).sum()) != pts.shape[0]:
        raise ValueError("sum(num_samples) must equal the number of points")
    if ns.size == 0:
        return np.zeros((0, int(target_num_samples), 3), np.float32)
    off = np.zeros(ns.size, np.int64)
    off[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    dev = torch.device("cuda")
    t_ns = torch.from_numpy(ns).to(dev)
    t_off = torch.from_numpy(off).to(dev)
    t_pts = torch.from_numpy(pts).to(dev)
    lens = torch.empty(pts.shape[0], dtype=torch.float32, device=dev)
    out = torch.empty((ns.size, int(target_num_samples), 3), dtype=torch.float32, device=dev)
    _core.resample_strands_by_arclength_device(t_ns, t_off, t_pts, int(target_num_samples), lens,
                                               out)
    return out.cpu().numpy()


def load_strands_arc_length(filepath: str, target_num_samples: int = 100) -> np.ndarray:
    """``load_strands`` then ``resample_strands_by_arclength`` (reference:
    __init__.py:97-102)."""
    _, ns, pts = load_strands(filepath, return_raw=True)
    return resample_strands_by_arclength(ns, pts, target_num_samples)


def _npz(npz_path):
    # numeric arrays only: no pickled objects are loaded
    with np.load(npz_path, allow_pickle=False) as data:
        return np.asarray(data["num_samples"]).ravel(), np.asarray(data["points"])


def load_npz_as_strands(npz_path) -> list[np.ndarray]:
    """List of (M_i, 3) strands from an .npz holding ``num_samples`` and
    ``points`` (reference: __init__.py:133-151)."""
    ns, pts = _npz(npz_path)
    ends = np.cumsum(ns.astype(np.int64))
    if ends.size and ends[-1] != len(pts) or (not ends.size and len(pts)):
        raise ValueError("Mismatch between num_samples and points length.")
    return [pts[e - n:e] for n, e in zip(ns.astype(np.int64), ends)]


def load_npz_as_strands_arc_length(npz_path, target_num_samples=100) -> np.ndarray:
    """(N, target, 3) resampled strands from such an .npz (reference:
    __init__.py:153-163)."""
    ns, pts = _npz(npz_path)
    return resample_strands_by_arclength(ns, pts, target_num_samples)


def resample_single_strand_by_arclength(strand: np.ndarray, target_samples: int) -> np.ndarray:
    """(target, 3) float32 linear resampling of one (N >= 2, 3) strand at
    evenly spaced arclengths (reference: __init__.py:177-204)."""
    strand = np.asarray(strand)
    if strand.shape[0] < 2:
        raise ValueError("Strand must have at least 2 points")
    s = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(strand, axis=0), axis=1))])
    u = np.linspace(0.0, s[-1], target_samples)
    return np.stack([np.interp(u, s, strand[:, d]) for d in range(3)], 1).astype(np.float32)


# ------------------------------------------------------------------ output


def save_obj(filename, verts, faces, uvs) -> None:
    """Writes an OBJ with per-vertex UVs (v flipped to 1 - v) and 1-based
    ``f a/a b/b c/c`` faces (reference: __init__.py:104-131)."""
    verts = np.asarray(verts)
    faces = np.asarray(faces)
    uvs = np.asarray(uvs)
    assert verts.shape[0] == uvs.shape[0], "Number of vertices and UVs must match."
    with open(filename, "w") as f:
        f.writelines(f"v {v[0]} {v[1]} {v[2]}\n" for v in verts)
        f.writelines(f"vt {t[0]} {1.0 - t[1]}\n" for t in uvs)
        for a, b, c in faces + 1:
            f.write(f"f {a}/{a} {b}/{b} {c}/{c}\n")
    print(f"Saved OBJ to {filename}")
