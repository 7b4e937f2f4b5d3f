"""Per-job card fitting on the CPU — TEST INFRASTRUCTURE AND CPU BASELINE ONLY.

The fitting iteration of 3_fitting.py:381-407 for one cluster and one view,
as the batched GPU kernel and the plain drop-in loop run it, composed from
the C oracle (triangle cluster mask, rasterize.cu / rasterize.py:170-214;
silhouette marking, DDA and antialias backward, antialias.py:55-130) and
float64 numpy for the rest:

* positions = M^{-1} param, M = I + lam (L_h + w L_w) of the strip
  (largesteps, 3_fitting.py:347-384), through a Cholesky factor;
* clip = mvp [pos, 1]; the card mask; the antialiased gradient w.r.t. clip
  (x, y, w), chained through mvp;
* planarity (3_fitting.py:181-199) and uniform-Laplacian
  (optimizer.py:70-99) gradients;
* param gradient = M^{-1} position gradient; UniformAdam (optimizer.py:33-68)
  with the exponential learning-rate decay.

The strip mesh and the strand edges of a job are built here as well
(strip_mesh, job_dict), so the loop depends on nothing in the product.
bench.py times it as the CPU path of card fitting (one process per core)
and the GPU tests compare the batched kernel with it; nothing in the
product calls it.
"""
from __future__ import annotations

import time

import numpy as np

from . import (antialiased_cluster_bitset_grad, dda_cluster_mask, laplacian_edges,
               largesteps_matrix, popcount_bitset, triangle_cluster_mask)


HYPER_FIELDS = ("lr", "beta1", "beta2", "gamma", "w_pln", "w_lap", "rho", "radius", "lam",
                "w_weight")


def strip_mesh(Q: int):
    """Two-row strip of Q quads as wisp_to_grid_mesh builds it
    (texture_utils.py:525-582 with S = 2, N = Q + 1): triangles [i0, i1, i2],
    [i0, i2, i3] and CCW quads [i0, i1, i2, i3] per grid cell, h_edges along
    each row, w_edges across; the unique edges with up to two adjacent
    triangles (build_edges_from_triangles, mesh_utils.cpp:36-80; listed in
    first-appearance order, which only changes summation order)."""
    N = Q + 1
    tri, quads = [], []
    for j in range(Q):
        i0, i1, i2, i3 = j, N + j, N + j + 1, j + 1
        tri += [[i0, i1, i2], [i0, i2, i3]]
        quads.append([i0, i1, i2, i3])
    h = [[r * N + j, r * N + j + 1] for r in range(2) for j in range(Q)]
    w = [[j, N + j] for j in range(N)]
    emap = {}
    for t, (a, b, c) in enumerate(tri):
        for u, v in ((a, b), (b, c), (c, a)):
            emap.setdefault((min(u, v), max(u, v)), []).append(t)
    edges = np.array(list(emap), np.int32)
    e2t = np.array([[ts[0], ts[1] if len(ts) > 1 else -1] for ts in emap.values()], np.int32)
    i32 = np.int32
    return (np.asarray(tri, i32), np.asarray(quads, i32), edges, e2t, np.asarray(h, i32),
            np.asarray(w, i32))


def job_dict(res: int, quads: int, strand_clip, mvp, verts) -> dict:
    """The fit_job input of one job: strand vertices in clip space (n, nv, 4),
    its consecutive-vertex strand edges (build_clustered_strand_edges,
    3_fitting.py:17-60), the strip mesh and the uniform-Laplacian edges
    h_edges + w_edges (3_fitting.py:317-319, 359)."""
    tri, qd, edges, e2t, h, w = strip_mesh(quads)
    clip = np.asarray(strand_clip, np.float32)
    n, nv = clip.shape[:2]
    base = np.arange(n * nv, dtype=np.int64).reshape(n, nv)
    se = np.stack([base[:, :-1].reshape(-1), base[:, 1:].reshape(-1)], 1).astype(np.int32)
    return dict(res=res, quads=quads, clip=clip.reshape(-1, 4), strand_edges=se,
                mvp=np.asarray(mvp, np.float32), verts=np.asarray(verts, np.float32), tri=tri,
                quad_idx=qd, edges=edges, edge2tri=e2t, lap_edges=np.concatenate([h, w]),
                h_edges=h, w_edges=w)


def hyper_dict(h) -> dict:
    """The fit_job hyper-parameters from any object with the FitHyper
    attributes."""
    return {k: float(getattr(h, k)) for k in HYPER_FIELDS}


def clip_grad_to_positions(gc, mvp):
    """Gradient w.r.t. positions of a loss whose gradient w.r.t. the clip
    coordinates clip = mvp [pos, 1] is gc (V, 4); the antialias gradient has
    no z component, so gc[:, 2] is ignored."""
    gc = np.asarray(gc, np.float64)
    mvp = np.asarray(mvp, np.float64)
    return gc[:, 0:1] * mvp[0, :3] + gc[:, 1:2] * mvp[1, :3] + gc[:, 3:4] * mvp[3, :3]


def regularizer_grad(pos, quads, L, w_pln, w_lap):
    """Gradient of w_pln * planarity + w_lap * |L pos|^2, with planarity =
    sum over quads of |(q0 + q2) - (q1 + q3)|^2 (3_fitting.py:181-199) and L
    the uniform Laplacian of the strip's h and w edges (optimizer.py:70-99)."""
    g = np.zeros_like(pos)
    d = (pos[quads[:, 0]] + pos[quads[:, 2]]) - (pos[quads[:, 1]] + pos[quads[:, 3]])
    for k, sgn in ((0, 2.0), (2, 2.0), (1, -2.0), (3, -2.0)):
        np.add.at(g, quads[:, k], sgn * w_pln * d)
    return g + 2.0 * w_lap * (L.T @ (L @ pos))


def fit_job(job: dict, hyper: dict, iters: int):
    """One job: `job` holds res, quads, clip (n*nv, 4) float32 strand
    vertices in the job's clip space, strand_edges (E_s, 2), mvp (4, 4),
    verts (2(Q+1), 3) and the strip topology tri, quad_idx, edges, edge2tri,
    lap_edges, h_edges, w_edges; `hyper` the FitHyper fields.  Returns
    (final vertices, final XOR pixel count, target pixel count)."""
    res, Q = int(job["res"]), int(job["quads"])
    V = 2 * (Q + 1)
    tri, quads, edges, e2t = job["tri"], job["quad_idx"], job["edges"], job["edge2tri"]
    zt = np.zeros(tri.shape[0], np.int32)
    ze = np.zeros(edges.shape[0], np.int32)
    clip_s = job["clip"].reshape(1, -1, 4)
    target = dda_cluster_mask((res, res), clip_s, job["strand_edges"],
                              np.zeros(job["strand_edges"].shape[0], np.int32), 1)
    mvp = np.asarray(job["mvp"], np.float64)
    L = laplacian_edges(V, job["lap_edges"])
    lam = float(hyper["lam"])
    if lam > 0:
        from scipy.linalg import cho_factor, cho_solve

        M = largesteps_matrix(V, job["h_edges"], job["w_edges"], lam, float(hyper["w_weight"]))
        fac = cho_factor(M)
        solve = lambda b: cho_solve(fac, b)  # noqa: E731
        param = M @ np.asarray(job["verts"], np.float64)
    else:
        solve = lambda b: b  # noqa: E731
        param = np.asarray(job["verts"], np.float64).copy()
    b1, b2 = float(hyper["beta1"]), float(hyper["beta2"])
    w_pln, w_lap = float(hyper["w_pln"]), float(hyper["w_lap"])
    lr, gamma = float(hyper["lr"]), float(hyper["gamma"])
    g1 = np.zeros((V, 3))
    g2 = np.zeros((V, 1))

    def clip_of(pos):
        return (np.concatenate([pos, np.ones((V, 1))], 1) @ mvp.T).astype(np.float32)

    for t in range(1, iters + 1):
        pos = solve(param)
        clip = clip_of(pos)
        render = triangle_cluster_mask((res, res), clip[None], tri, zt, 1)
        gc = antialiased_cluster_bitset_grad(render, target, clip[None], tri, edges, e2t, ze,
                                             float(hyper["radius"]), float(hyper["rho"]))[0]
        g = clip_grad_to_positions(gc, mvp) + regularizer_grad(pos, quads, L, w_pln, w_lap)
        g = solve(g)
        g1 = b1 * g1 + (1.0 - b1) * g
        g2 = b2 * g2 + (1.0 - b2) * (g * g).sum(1, keepdims=True)
        denom = np.sqrt(g2 / (1.0 - b2 ** t)).max() + 1e-8
        param = param - lr * (g1 / (1.0 - b1 ** t)) / denom
        lr *= gamma
    pos = solve(param)
    render = triangle_cluster_mask((res, res), clip_of(pos)[None], tri, zt, 1)
    xor = int(popcount_bitset(np.bitwise_xor(render, target)).sum())
    return pos.astype(np.float32), xor, int(popcount_bitset(target).sum())


def fit_jobs(args):
    """Pool worker: fits a list of jobs in this process (one thread) and
    returns (iterations, seconds, sum of XOR counts, sum of target pixels)."""
    jobs, hyper, iters = args
    t0 = time.perf_counter()
    xor = tp = 0
    for job in jobs:
        _, x, p = fit_job(job, hyper, iters)
        xor += x
        tp += p
    return len(jobs) * iters, time.perf_counter() - t0, xor, tp
