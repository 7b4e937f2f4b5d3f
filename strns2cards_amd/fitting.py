"""Per-cluster card fitting: the plain per-job loop and the batched / served
form.

A fitting job is one cluster (wisp) of strands and one card strip fitted to
it in one orthographic view, with the iteration of 3_fitting.py:381-407:
triangle cluster mask, antialiased silhouette gradient (rho, radius 7),
planarity and uniform-Laplacian regularizers, UniformAdam and an
exponential learning-rate decay, over the largesteps parameterization of
3_fitting.py:347-384 (param = M pos, pos = M^{-1} param, M = I + lam (L_h +
w L_w)) in both forms: ``optimizer.GridSystem`` in the plain loop, the same
factor and solve inside the batched kernel.

* ``plain_fit`` runs one job at a time through the drop-in API
  (``compute_fragments`` -> ``cluster_mask_from_fragments`` ->
  ``antialiased_cluster_bitset_loss`` -> autograd -> ``UniformAdam``), with
  the host synchronizations that API implies.
* ``CardFitServer`` keeps every admitted job's state resident in HBM (slots)
  and advances all active slots with one ``fit_cards`` launch (one
  workgroup per job, masks in LDS); ``serve`` feeds it with arriving jobs
  (continuous batching) and replays a captured HIP graph per chunk.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _core
from . import synth
from .antialias import antialiased_cluster_bitset_loss
from .bitset_utils import popcount_bitset
from .optimizer import (GridSystem, UniformAdam, from_differential, grid_laplacian, grid_quads,
                        grid_quad_diagonals, laplacian_edges, to_differential)
from .rasterize import cluster_mask_from_fragments, compute_fragments, triangle_cluster_mask

MAX_RES = 256
MAX_QUADS = 32
MAX_VERTS = 2 * (MAX_QUADS + 1)
MAX_WORDS = MAX_RES * MAX_RES // 32
# per-slot scalars (res, quads, offsets, steps, iteration limit, loss, slot
# ids) and the 4 x 4 clip matrix
SLOT_SCALAR_BYTES = 9 * 4 + 16 * 4
NO_LIMIT = 2**31 - 1


@dataclass
class FitHyper:
    """3_fitting.py:506-525 defaults (lr 0.1, alpha 0.5, w_pln 10,
    w_lap 1e-2, kernel radius 7, laplacian_lambda 100, laplacian_w_weight
    1/100) with the decay to 1e-2 of the initial lr over the job's
    iterations (3_fitting.py:376).  lam = 0 optimizes the positions
    directly (no largesteps)."""
    lr: float = 0.1
    beta1: float = 0.9
    beta2: float = 0.999
    iters: int = 100
    w_pln: float = 10.0
    w_lap: float = 1e-2
    rho: float = 0.5
    radius: float = 7.0
    lam: float = 100.0
    w_weight: float = 1.0 / 100

    @property
    def gamma(self) -> float:
        return 1e-2 ** (1.0 / self.iters)

    def vector(self) -> list[float]:
        return [self.lr, self.beta1, self.beta2, self.gamma, self.w_pln, self.w_lap, self.rho,
                self.radius, self.lam, self.w_weight]


@dataclass
class FitJob:
    jid: int
    res: int
    quads: int
    strands: np.ndarray        # (n, nv, 3) cm
    strand_clip: np.ndarray    # (n * nv, 4) float32, clip space of the job's view
    mvp: np.ndarray            # (4, 4) float32
    verts: np.ndarray          # (2 (Q + 1), 3) float32 initial strip
    arrival: float = 0.0

    @property
    def words(self) -> int:
        """32-bit words of the packed res x res target mask."""
        return (self.res * self.res + 31) // 32

    @property
    def bytes(self) -> int:
        """HBM footprint while resident: the target mask, the vertices, the
        large-steps parameter and the Adam moments (40 B per vertex), and
        the slot's per-job scalars and clip matrix.  The strand vertices are
        uploaded for the target computation at admission and released."""
        return 4 * self.words + 40 * self.verts.shape[0] + SLOT_SCALAR_BYTES

    def validate(self) -> None:
        """Raises ValueError for a job the fitting kernel cannot hold."""
        if not (1 <= self.res <= MAX_RES):
            raise ValueError(f"job {self.jid}: resolution {self.res} outside [1, {MAX_RES}]")
        if not (1 <= self.quads <= MAX_QUADS):
            raise ValueError(f"job {self.jid}: {self.quads} quads outside [1, {MAX_QUADS}]")
        if self.verts.shape != (2 * (self.quads + 1), 3):
            raise ValueError(f"job {self.jid}: the strip needs {2 * (self.quads + 1)} vertices")
        if self.strands.ndim != 3 or self.strands.shape[1] < 2:
            raise ValueError(f"job {self.jid}: strands must be (n, nv >= 2, 3)")


def make_jobs(count: int, seed: int = 0, nv: int = 64, cap_angle: float = 0.15) -> list[FitJob]:
    """Seeded per-cluster jobs: strand count log-uniform in [8, 512], nv
    vertices each (step 0.004, strand model of synth.generate_strands, roots
    in a cap of `cap_angle` rad around a random direction on the upper
    hemisphere, scaled to cm as 3_fitting.py:289 does), quads uniform in
    {8, 12, 16, 24, 32}, resolution uniform in {64, 128, 256}.  The card is
    the mean strand resampled to Q + 1 points, swept 1 cm wide."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(8), math.log(512), size=count))).astype(np.int64)
    Qs = rng.choice([8, 12, 16, 24, 32], size=count)
    ress = rng.choice([64, 128, 256], size=count)
    centres = rng.normal(size=(count, 3))
    centres[:, 1] = np.abs(centres[:, 1])
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    owner = np.repeat(np.arange(count), ns)
    d = centres[owner] + cap_angle * rng.normal(size=(owner.size, 3))
    d[:, 1] = np.abs(d[:, 1])
    # all jobs' strands in one vectorized pass, then split per job
    strands_all = synth.generate_strands(owner.size, nv, 0.004, seed, rng=rng, root_dirs=d) * 100.0
    starts = np.concatenate([[0], np.cumsum(ns)])
    jobs = []
    for j in range(count):
        strands = strands_all[starts[j]:starts[j + 1]]
        pts = strands.reshape(-1, 3)
        mvp = synth.ortho_cameras(pts, 1, 0.3)[0]
        clip = synth.project(pts, mvp[None])[0]
        jobs.append(FitJob(j, int(ress[j]), int(Qs[j]), strands.astype(np.float32), clip,
                           mvp.astype(np.float32), np.zeros((2 * (int(Qs[j]) + 1), 3), np.float32)))
    # cards: mean strand -> Q + 1 points -> 1 cm strip, batched per quad count
    for Q in np.unique(Qs):
        sel = [j for j in range(count) if Qs[j] == Q]
        means = np.stack([jobs[j].strands.mean(0) for j in sel])
        rows = synth.strip_rows(synth.resample_by_arclength(means, int(Q) + 1), 1.0)
        for k, j in enumerate(sel):
            jobs[j].verts = rows[k].reshape(-1, 3).astype(np.float32)
    return jobs


# ----------------------------------------------------------------- topology


def strip_topology(Q: int):
    """Triangles, quads, unique edges (+ adjacency) and Laplacian edges of a
    two-row strip with Q quads (wisp_to_grid_mesh, texture_utils.py:525-599)."""
    N = Q + 1
    tri = synth.grid_triangles(2, N)
    quads = np.array([[c, N + c, N + c + 1, c + 1] for c in range(Q)], np.int32)
    edges, e2t = synth.edges_from_triangles(tri)
    h_edges = np.array([[c, c + 1] for c in range(Q)] + [[N + c, N + c + 1] for c in range(Q)],
                       np.int32)
    w_edges = np.array([[c, N + c] for c in range(N)], np.int32)
    return tri, quads, edges, e2t, np.concatenate([h_edges, w_edges])


def planarity_loss(pos: torch.Tensor, quads: torch.Tensor) -> torch.Tensor:
    """sum |(q0 + q2) - (q1 + q3)|^2 over quads (3_fitting.py:181-199)."""
    d = (pos[quads[:, 0]] + pos[quads[:, 2]]) - (pos[quads[:, 1]] + pos[quads[:, 3]])
    return (d ** 2).sum()


# --------------------------------------------------------------- plain loop


def render_mask(res: int, pos_ndc: torch.Tensor, tri: torch.Tensor) -> torch.Tensor:
    """render_bitset_from_triangles (3_fitting.py:154-163) for one cluster;
    an empty render (card out of view) is an all-zero mask instead of the
    reference's empty-max error."""
    fp, fa = compute_fragments((res, res), pos_ndc, tri)
    if fp.shape[0] == 0:
        return torch.zeros((pos_ndc.shape[0], res, res, 1), dtype=torch.int32, device=pos_ndc.device)
    t2c = torch.zeros(tri.shape[0], dtype=torch.int32, device=tri.device)
    bits = cluster_mask_from_fragments((res, res), fp, fa, t2c, cluster_count=1)
    if bits.shape[0] < pos_ndc.shape[0]:
        pad = torch.zeros((pos_ndc.shape[0] - bits.shape[0], res, res, 1), dtype=bits.dtype,
                          device=bits.device)
        bits = torch.cat([bits, pad])
    return bits


def plain_target(job: FitJob, dev) -> torch.Tensor:
    """Target mask through the drop-in API: dda_compute_fragments +
    cluster_mask_from_fragments with every edge in cluster 0."""
    from .dda import dda_compute_fragments

    n, nv = job.strands.shape[:2]
    pos = torch.from_numpy(job.strand_clip[None]).to(dev)
    edges = torch.from_numpy(synth.strand_edges(n, nv)).to(dev)
    fp, fa = dda_compute_fragments((job.res, job.res), pos, edges)
    if fp.shape[0] == 0:
        return torch.zeros((1, job.res, job.res, 1), dtype=torch.int32, device=dev)
    e2c = torch.zeros(edges.shape[0], dtype=torch.int32, device=dev)
    return cluster_mask_from_fragments((job.res, job.res), fp, fa, e2c, cluster_count=1)


def plain_fit(job: FitJob, hyper: FitHyper, dev, iters: int | None = None):
    """One job through the drop-in API, as 3_fitting.py:362-407: param =
    to_differential(M, pos), and each iteration pos = from_differential(M,
    param) (``GridSystem`` of the strip; positions directly when lam = 0).
    Returns (final vertices, final XOR pixel count, target pixel count)."""
    iters = hyper.iters if iters is None else iters
    tri_np, quads_np, edges_np, e2t_np, lap_np = strip_topology(job.quads)
    tri = torch.from_numpy(tri_np).to(dev)
    quads = torch.from_numpy(quads_np).to(dev)
    edges = torch.from_numpy(edges_np).to(dev)
    e2t = torch.from_numpy(e2t_np).to(dev)
    e2c = torch.zeros(edges.shape[0], dtype=torch.int32, device=dev)
    mvp = torch.from_numpy(job.mvp).to(dev)
    target = plain_target(job, dev)
    pos0 = torch.from_numpy(job.verts.copy()).to(dev)
    L = laplacian_edges(pos0, torch.from_numpy(lap_np).to(dev))
    if hyper.lam > 0:
        M = GridSystem([2], [job.quads + 1], hyper.lam, hyper.w_weight, device=dev)
        param = to_differential(M, pos0).detach().requires_grad_(True)
        positions = lambda: from_differential(M, param)  # noqa: E731
    else:
        param = pos0.clone().requires_grad_(True)
        positions = lambda: param  # noqa: E731
    opt = UniformAdam([param], lr=hyper.lr, betas=(hyper.beta1, hyper.beta2))
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=hyper.gamma)

    def project(p):
        ph = torch.cat([p, torch.ones_like(p[..., :1])], -1)
        return torch.matmul(ph, mvp.t())[None].contiguous()

    for _ in range(iters):
        opt.zero_grad()
        pos = positions()
        pos_ndc = project(pos)
        bits = render_mask(job.res, pos_ndc.detach(), tri)
        loss = antialiased_cluster_bitset_loss(bits, target, pos_ndc, tri, edges, e2t, e2c,
                                               kernel_radius=hyper.radius, rho=hyper.rho)
        loss = loss + hyper.w_pln * planarity_loss(pos, quads)
        loss = loss + hyper.w_lap * torch.sum(torch.sparse.mm(L, pos) ** 2)
        loss.backward()
        opt.step()
        sched.step()
    with torch.no_grad():
        pos = positions()
        bits = render_mask(job.res, project(pos), tri)
        xor = int(popcount_bitset(torch.bitwise_xor(bits, target)).sum().item())
        tpix = int(popcount_bitset(target).sum().item())
    return pos.detach().cpu().numpy(), xor, tpix


# ------------------------------------------------------------- batched form


@dataclass
class SlotState:
    job: FitJob
    admitted: float
    done_iters: int = 0


class WordPool:
    """Allocator of target words in one device buffer: blocks have
    power-of-two sizes, a freed block returns to the free list of its size,
    and a bump pointer serves a size with no free block.  ``alloc`` returns
    None when the buffer cannot hold the block."""

    def __init__(self, words: int):
        self.words = int(words)
        self.top = 0
        self.free: dict[int, list[int]] = {}

    @staticmethod
    def block(words: int) -> int:
        return 1 << max(0, int(words - 1).bit_length())

    def can_alloc(self, words: int) -> bool:
        b = self.block(words)
        return bool(self.free.get(b)) or self.top + b <= self.words

    def alloc(self, words: int):
        b = self.block(words)
        if self.free.get(b):
            return self.free[b].pop()
        if self.top + b > self.words:
            return None
        off = self.top
        self.top += b
        return off

    def release(self, off: int, words: int) -> None:
        self.free.setdefault(self.block(words), []).append(off)


def _upload(a, dev, dtype=None) -> torch.Tensor:
    """Host array -> device tensor through pinned memory, queued on the
    current stream without waiting for the work already queued on it.  The
    pinned block comes from torch's caching host allocator (reused once
    the copy has completed) and is filled by numpy: a torch host copy
    would start its thread pool, whose spinning threads cost a serving
    loop its CPU share."""
    a = np.asarray(a if dtype is None else np.asarray(a, dtype))
    buf = torch.empty(a.shape, dtype=torch.from_numpy(a[:0].reshape(-1)).dtype, pin_memory=True)
    buf.numpy()[...] = a
    return buf.to(dev, non_blocking=True)


class CardFitServer:
    """Device-resident state of fitting jobs, admitted by HBM budget.

    Slot s (one workgroup of the ``fit_cards`` launch) holds a job's per-job
    scalars, clip matrix and its vertex state at rows [s*MAX_VERTS, ...);
    the job's target mask takes a block of a word pool sized from the
    budget.  The budget pays for the slot tables first; what is left (up
    to what ``capacity`` masks at the largest resolution need) is the pool.
    A job is admitted when a slot is free and the pool can hold its mask;
    otherwise it waits.  ``run(iters)`` advances every active slot
    with one ``fit_cards`` launch whose grid is the whole table (idle slots
    exit at once), so the launch can be captured once and replayed while
    jobs come and go."""

    def __init__(self, capacity: int, hyper: FitHyper, dev, budget_bytes: float = 200e9,
                 chunk: int = 0, order: str = "admission"):
        """``capacity``: the most jobs resident at once (the launch grid).
        ``budget_bytes``: HBM for the jobs' resident state; the slot tables
        take their part first, and the target pool gets the rest, up to
        what ``capacity`` jobs at the largest resolution would need.
        ``chunk`` (silhouette samples per workgroup pass: 0 = 512, 256,
        512, 1024) and ``order`` ("admission": slot order; "cost": jobs
        with the most work first) are launch knobs chosen by
        ``autotune.tune_fit``; no job's result depends on them."""
        if order not in ("admission", "cost"):
            raise ValueError("order must be 'admission' or 'cost'")
        table = capacity * (40 * MAX_VERTS + SLOT_SCALAR_BYTES)
        if budget_bytes < table + 4 * MAX_WORDS:
            raise ValueError(f"budget of {budget_bytes:.0f} B cannot hold {capacity} slots "
                             f"({table} B) and one {MAX_RES}^2 mask")
        self.cap = capacity
        self.chunk = int(chunk)
        self.order = order
        self.hyper = hyper
        self.dev = dev
        self.budget = float(budget_bytes)
        self.used_bytes = 0
        pool_words = int(min((budget_bytes - table) // 4, capacity * MAX_WORDS))
        self.pool = WordPool(pool_words)
        i32 = dict(dtype=torch.int32, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        self.res = torch.zeros(capacity, **i32)
        self.quads = torch.ones(capacity, **i32)
        self.vert_off = torch.arange(capacity, **i32) * MAX_VERTS
        self.bits_off = torch.zeros(capacity, **i32)
        self.mvp = torch.zeros((capacity, 4, 4), **f32)
        self.steps = torch.zeros(capacity, **i32)
        self.loss = torch.zeros(capacity, **f32)
        self.verts = torch.zeros((capacity * MAX_VERTS, 3), **f32)
        self.params = torch.zeros((capacity * MAX_VERTS, 3), **f32)
        self.adam_m = torch.zeros((capacity * MAX_VERTS, 3), **f32)
        self.adam_v = torch.zeros(capacity * MAX_VERTS, **f32)
        self.target = torch.zeros(max(pool_words, 1), **i32)
        self.limit = torch.full((capacity,), NO_LIMIT, **i32)
        self.slots = torch.full((capacity,), -1, **i32)
        self.launch = torch.full((capacity,), -1, **i32)  # workgroup -> slot ("cost" order)
        self.launch_dirty = False
        self.free = deque(range(capacity))
        self.state: dict[int, SlotState] = {}
        self.word_off: dict[int, int] = {}
        self.graphs: dict[tuple[int, bool], torch.cuda.CUDAGraph] = {}

    # -- admission -----------------------------------------------------
    def can_admit(self, job: FitJob, pending_words=()) -> bool:
        """True when `job` fits now: a free slot and room in the target pool
        (beyond the masks in `pending_words`, promised to jobs not yet
        admitted).  Raises ValueError for a job the kernel cannot hold."""
        job.validate()
        if len(self.free) <= len(pending_words):
            return False
        if not pending_words:
            return self.pool.can_alloc(job.words)
        trial = WordPool(self.pool.words)
        trial.top, trial.free = self.pool.top, {k: list(v) for k, v in self.pool.free.items()}
        for w in pending_words:
            if trial.alloc(w) is None:
                return False
        return trial.can_alloc(job.words)

    def admit(self, jobs: list[FitJob], now: float = 0.0, limit: int | None = None) -> list[int]:
        """Uploads the jobs into free slots, computes their target masks on
        the GPU and activates the slots (all asynchronous on the current
        stream: one packed host-to-device copy of integers and one of
        floats, then the admission and target kernels).  With `limit`, each
        job stops after that many iterations however many later launches
        ask for.  Raises ValueError if one does not fit (``can_admit``)."""
        if not jobs:
            return []
        for j in jobs:
            j.validate()
        if len(jobs) > len(self.free):
            raise ValueError(f"{len(jobs)} jobs for {len(self.free)} free slots")
        nv = jobs[0].strands.shape[1]
        if any(j.strands.shape[1] != nv for j in jobs):
            raise ValueError("jobs admitted together need the same points per strand")
        if limit is not None and not 0 <= limit <= NO_LIMIT:
            raise ValueError("limit must be in [0, 2^31 - 1]")
        offs_w = []
        for j in jobs:
            o = self.pool.alloc(j.words)
            if o is None:
                for k, jj in zip(offs_w, jobs):
                    self.pool.release(k, jj.words)
                raise ValueError(f"job {j.jid} does not fit the HBM budget")
            offs_w.append(o)
        self.used_bytes += sum(j.bytes for j in jobs)
        slots = [self.free.popleft() for _ in jobs]
        n = len(jobs)
        counts = np.array([j.strands.shape[0] for j in jobs], np.int64)
        offs = np.concatenate([[0], np.cumsum(counts)[:-1]]) * nv
        # integers: n admission records (slot, res, quads, word offset,
        # vertex count, limit, 0, 0), then the target kernel's per-job
        # res, word offset, strand offset and strand count
        ib = np.zeros(12 * n, np.int32)
        rec = ib[: 8 * n].reshape(n, 8)
        rec[:, 0] = slots
        rec[:, 1] = [j.res for j in jobs]
        rec[:, 2] = [j.quads for j in jobs]
        rec[:, 3] = offs_w
        rec[:, 4] = [j.verts.shape[0] for j in jobs]
        rec[:, 5] = NO_LIMIT if limit is None else limit
        tail = ib[8 * n:].reshape(4, n)
        tail[0], tail[1], tail[2], tail[3] = rec[:, 1], rec[:, 3], offs, counts
        # floats: n records (clip matrix, vertices), then the strand vertices
        F = 16 + 3 * MAX_VERTS
        fb = np.zeros(n * F + int(counts.sum()) * nv * 4, np.float32)
        fr = fb[: n * F].reshape(n, F)
        for k, j in enumerate(jobs):
            fr[k, :16] = j.mvp.reshape(-1)
            fr[k, 16: 16 + j.verts.size] = j.verts.reshape(-1)
        fb[n * F:] = np.concatenate([j.strand_clip.reshape(-1) for j in jobs])
        idev, fdev = _upload(ib, self.dev), _upload(fb, self.dev)
        _core.fit_admit(idev[: 8 * n].view(n, 8), fdev[: n * F].view(n, F), self.slots, self.res,
                        self.quads, self.bits_off, self.limit, self.mvp, self.steps, self.verts,
                        self.adam_m, self.adam_v)
        t = idev[8 * n:].view(4, n)
        _core.fit_targets(t[0], t[1], t[2], t[3], fdev[n * F:].view(-1, nv, 4), self.target)
        for s, j, o in zip(slots, jobs, offs_w):
            self.state[s] = SlotState(j, now)
            self.word_off[s] = o
        self.launch_dirty = True
        return slots

    @staticmethod
    def job_cost(job: FitJob) -> float:
        """Relative work of one iteration of a job: mask clears / loss over
        res^2 bits and silhouette samples along ~res-long edges."""
        return job.res * job.res / 32.0 + 4.0 * job.res * job.quads

    def _refresh_launch(self) -> None:
        act = sorted(self.state, key=lambda s: -self.job_cost(self.state[s].job))
        arr = np.full(self.cap, -1, np.int32)
        arr[: len(act)] = act
        self.launch.copy_(_upload(arr, self.dev))
        self.launch_dirty = False

    def _forget(self, slots: list[int]) -> None:
        """Host side of a release: the slots and their budget return to the
        free lists.  The device tables are left as they are (a job that has
        reached its iteration limit is not advanced, and the next admission
        into the slot overwrites them)."""
        for s in slots:
            job = self.state.pop(s).job
            self.used_bytes -= job.bytes
            self.pool.release(self.word_off.pop(s), job.words)
            self.free.append(s)
        self.launch_dirty = True

    def release(self, slots: list[int]) -> None:
        """Frees the slots and their budget (stream-ordered: launches
        already queued still see them)."""
        if not slots:
            return
        self.slots[_upload(slots, self.dev, np.int64)] = -1
        self._forget(slots)

    # -- iterations ----------------------------------------------------
    def _launch(self, iters: int, order: torch.Tensor) -> None:
        _core.fit_cards(order, iters, self.res, self.quads, self.vert_off, self.bits_off,
                        self.mvp, self.verts, self.params, self.adam_m, self.adam_v, self.steps,
                        self.limit, self.target, self.hyper.vector(), self.loss, self.chunk)

    def run(self, iters: int, graph: bool = False, ordered: bool = True) -> None:
        """Advances every active slot by `iters` iterations, or by what is
        left of its limit (and refreshes its loss).  With graph=True the
        launch is captured once per iteration count and replayed.  With
        ordered=False the launch walks the slot table instead of the
        "cost" order (no host work when jobs come and go; what a serving
        loop with few resident jobs wants)."""
        cost = ordered and self.order == "cost"
        if cost and self.launch_dirty:
            self._refresh_launch()
        order = self.launch if cost else self.slots
        if not graph:
            self._launch(iters, order)
        else:
            g = self.graphs.get((iters, cost))
            if g is None:
                s = torch.cuda.Stream()
                s.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(s):
                    self._launch(0, order)  # warm-up outside capture (iters = 0 writes only the loss)
                torch.cuda.current_stream().wait_stream(s)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._launch(iters, order)
                self.graphs[(iters, cost)] = g
            g.replay()
        for st in self.state.values():
            st.done_iters += iters

    def job_vertices(self, slot: int) -> np.ndarray:
        n = self.state[slot].job.verts.shape[0]
        return self.verts[slot * MAX_VERTS: slot * MAX_VERTS + n].cpu().numpy()

    def target_mask(self, slot: int) -> np.ndarray:
        """(res, res) bool target of a slot, unpacked from its bit words."""
        job = self.state[slot].job
        res, o = job.res, self.word_off[slot]
        words = self.target[o: o + job.words].cpu().numpy()
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return bits[: res * res].reshape(res, res).astype(bool)


@dataclass
class ServeReport:
    completed: int
    iterations: int
    seconds: float
    latencies_ms: list[float] = field(default_factory=list)
    losses: dict[int, float] = field(default_factory=dict)
    vertices: dict[int, np.ndarray] = field(default_factory=dict)

    @property
    def iters_per_s(self) -> float:
        return self.iterations / self.seconds if self.seconds > 0 else 0.0

    def percentile(self, q: float) -> float:
        return float(np.percentile(self.latencies_ms, q)) if self.latencies_ms else float("nan")


def serve(jobs: list[FitJob], server: CardFitServer, chunk: int = 10, graph: bool = True,
          clock=time.perf_counter, keep_vertices: bool = False, ahead: int = 1) -> ServeReport:
    """Continuous batching: jobs (with `arrival` seconds) are admitted as
    slots and budget free up, and all resident jobs advance together
    `chunk` iterations per replay (in slot order; each job is admitted with
    its iteration count as its limit, so a finished job is not advanced and
    its slot is reused without any device work).  The host does not wait
    for a replay before queueing the next (at most `ahead` are queued
    beyond the one running): after a replay that finishes jobs, the loss
    table (and, with `keep_vertices`, the finished jobs' vertices) is
    copied to pinned host memory behind it, and the results are collected,
    with each job's latency (arrival -> result on the host), once that
    copy's event has completed, while arrivals are admitted meanwhile."""
    total = server.hyper.iters
    if total % chunk:
        raise ValueError("chunk must divide the per-job iteration count")
    dev = server.dev
    pending = deque(sorted(jobs, key=lambda j: j.arrival))
    rep = ServeReport(0, 0, 0.0)
    results = deque()  # (event, [(slot, jid, arrival, n_verts)], pinned loss table, vertices)
    queued = deque()   # one event per queued replay
    # pinned loss tables free for reuse (more are allocated if results pile up)
    tables = [torch.empty(server.cap, dtype=torch.float32, pin_memory=True)
              for _ in range(ahead + 2)]
    stream = torch.cuda.current_stream(dev)
    if graph and not server.state:
        server.run(chunk, graph=True, ordered=False)  # capture before the clock starts
        stream.synchronize()
    t0 = clock()

    def collect(wait: bool) -> None:
        while results and (wait or results[0][0].query()):
            ev, done, lh, vh = results.popleft()
            ev.synchronize()
            t_done = clock() - t0
            for k, (slot, jid, arrival, n) in enumerate(done):
                rep.latencies_ms.append((t_done - arrival) * 1e3)
                rep.losses[jid] = float(lh[slot])
                if vh is not None:
                    rep.vertices[jid] = vh[k, :n].numpy().copy()
                rep.completed += 1
                rep.iterations += total
            tables.append(lh)
            wait = False

    while pending or server.state or results:
        now = clock() - t0
        batch, pw = [], []
        while pending and pending[0].arrival <= now and server.can_admit(pending[0], pw):
            j = pending.popleft()
            batch.append(j)
            pw.append(j.words)
        if batch:
            server.admit(batch, now, limit=total)
        if not server.state:
            if results:
                collect(wait=True)
            elif pending:
                time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
            continue
        server.run(chunk, graph=graph, ordered=False)
        ev = torch.cuda.Event()
        ev.record(stream)
        queued.append(ev)
        done = [sl for sl, st in server.state.items() if st.done_iters >= total]
        if done:
            lh = tables.pop() if tables else torch.empty(server.cap, dtype=torch.float32,
                                                         pin_memory=True)
            lh.copy_(server.loss, non_blocking=True)
            vh = None
            if keep_vertices:
                idx = _upload(done, dev, np.int64)
                rows = (idx[:, None] * MAX_VERTS + torch.arange(MAX_VERTS, device=dev)).reshape(-1)
                vh = torch.empty((len(done), MAX_VERTS, 3), dtype=torch.float32, pin_memory=True)
                vh.copy_(server.verts[rows].reshape(len(done), MAX_VERTS, 3), non_blocking=True)
            ev_done = torch.cuda.Event()
            ev_done.record(stream)
            results.append((ev_done, [(sl, server.state[sl].job.jid, server.state[sl].job.arrival,
                                       server.state[sl].job.verts.shape[0]) for sl in done],
                            lh, vh))
            server._forget(done)
        while len(queued) > ahead:
            queued.popleft().synchronize()
        collect(wait=False)
    rep.seconds = clock() - t0
    return rep


def poisson_arrivals(jobs: list[FitJob], rate: float, seed: int = 0) -> list[FitJob]:
    """Assigns seeded exponential inter-arrival times (mean 1/rate s)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=len(jobs)))
    for j, a in zip(jobs, t):
        j.arrival = float(a)
    return jobs


# ------------------------------------------------------------ joint fit


@dataclass
class JointFitReport:
    pos: torch.Tensor          # (V, 3) fitted card vertices
    xor_first: int             # popcount(render XOR target) over all views, before
    xor_last: int              # ... and after the iterations
    ms_per_iter: float


def joint_fit(points: torch.Tensor, edges: torch.Tensor, edge2cluster: torch.Tensor,
              cluster_count: int, card_pos: torch.Tensor, tri: torch.Tensor,
              tri2cluster: torch.Tensor, quads: torch.Tensor, h_edges: torch.Tensor,
              w_edges: torch.Tensor, resolution: int = 200, num_iters: int = 1000,
              lr: float = 0.1, alpha: float = 0.5, lam: float = 100.0, w_weight: float = 1 / 100,
              w_pln: float = 10.0, w_lap: float = 1e-2, margin: float = 0.3,
              mvp: torch.Tensor | None = None, fused: bool = False) -> JointFitReport:
    """The fitting loop of 3_fitting.py:232-407 (all cards, all views) on
    the drop-in API, with its defaults: 13 orthographic views
    (build_ortho_cameras), target cluster masks from the DDA of every strand
    edge, and per iteration pos = from_differential(M, param), the card
    cluster masks (render_bitset_from_triangles), the antialiased loss with
    kernel radius 7 and rho = alpha, planarity and uniform-Laplacian terms,
    UniformAdam and an exponential lr decay to 1e-2 over num_iters.  M is
    the card grids' GridSystem (lam, w_weight).  Device tensors: points
    (N, 3), edges (E, 2), edge2cluster (E,), card_pos (V, 3), tri (T, 3),
    tri2cluster (T,), quads (Q, 4), h_edges, w_edges.  `fused` renders the
    card masks with triangle_cluster_mask (the same masks in one pass, no
    fragment list or host sync) instead of the script's compute_fragments +
    cluster_mask_from_fragments, applies the uniform Laplacian of the w_lap
    term with the grid kernel (grid_laplacian) instead of a sparse matrix
    product, whose backward re-sorts the matrix every iteration, and forms
    the planarity term with grid_quad_diagonals instead of four gathers,
    whose backward scatters sort their indices, and takes the silhouette
    gradient from the marked edges in one launch sequence
    (antialiased_cluster_bitset_loss(fused=True)), so the iteration makes
    no host synchronization, and takes the UniformAdam step in two launches
    (UniformAdam(fused=True)).  The fused planarity term differentiates
    the grids' quads in wisp_to_grid_mesh order (as combine_wisps_with_
    cluster_labels writes them); if `quads` is anything else it falls back
    to planarity_loss(pos, quads), so the loss stays 3_fitting.py's."""
    from . import strand_io
    from .dda import dda_compute_fragments

    dev = points.device
    H = W = int(resolution)
    if mvp is None:
        mvp = torch.from_numpy(synth.ortho_cameras(points.detach().cpu().numpy(), 13, margin))
    mvp = mvp.to(dev, torch.float32)
    ph = torch.cat([points, torch.ones_like(points[:, :1])], -1)[None]
    pts_ndc = torch.matmul(ph, mvp.transpose(1, 2)).contiguous()
    fp, fa = dda_compute_fragments((H, W), pts_ndc, edges)
    target = cluster_mask_from_fragments((H, W), fp, fa, edge2cluster, cluster_count=cluster_count)

    cedges_np, e2t_np = strand_io.build_edges_from_triangles(tri.cpu().numpy().astype(np.int32))
    cedges = torch.from_numpy(cedges_np).to(dev)
    e2t = torch.from_numpy(e2t_np).to(dev)
    ce2c = tri2cluster[e2t[:, 0].long()].contiguous()
    V = card_pos.shape[0]
    M = GridSystem.from_edges(V, h_edges, w_edges, lam, w_weight, device=dev)
    L_uniform = None if fused else laplacian_edges(card_pos, torch.cat([h_edges, w_edges], 0))
    grid_pln = fused and np.array_equal(quads.detach().cpu().numpy().astype(np.int64),
                                        grid_quads(M.rows_host, M.cols_host))
    param = to_differential(M, card_pos.float()).detach().requires_grad_(True)
    opt = UniformAdam([param], lr=lr, betas=(0.9, 0.999), fused=fused)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=1e-2 ** (1 / num_iters))
    mt = mvp.transpose(1, 2)

    def render(pos):
        pos_h = torch.cat([pos, torch.ones_like(pos[..., :1])], -1)
        pos_ndc_h = torch.matmul(pos_h, mt).contiguous()
        if fused:
            return pos_ndc_h, triangle_cluster_mask((H, W), pos_ndc_h.detach(), tri, tri2cluster,
                                                    cluster_count)
        f_pix, f_attr = compute_fragments((H, W), pos_ndc_h.detach(), tri)
        bits = cluster_mask_from_fragments((H, W), f_pix, f_attr, tri2cluster,
                                           cluster_count=cluster_count)
        return pos_ndc_h, bits

    def xor_count(bits):
        return int(popcount_bitset(torch.bitwise_xor(bits, target)).sum().item())

    with torch.no_grad():
        xor_first = xor_count(render(from_differential(M, param))[1])
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(num_iters):
        opt.zero_grad()
        pos = from_differential(M, param)
        pos_ndc_h, bits = render(pos)
        loss = antialiased_cluster_bitset_loss(bits, target, pos_ndc_h, tri, cedges, e2t, ce2c,
                                               kernel_radius=7.0, rho=alpha, fused=fused)
        if grid_pln:
            loss = loss + w_pln * torch.sum(grid_quad_diagonals(M, pos) ** 2)
        else:
            loss = loss + w_pln * planarity_loss(pos, quads)
        lap = grid_laplacian(M, pos) if fused else torch.sparse.mm(L_uniform, pos)
        loss = loss + w_lap * torch.sum(lap ** 2)
        loss.backward()
        opt.step()
        sched.step()
    torch.cuda.synchronize(dev)
    ms = (time.perf_counter() - t0) * 1e3 / max(num_iters, 1)
    with torch.no_grad():
        pos = from_differential(M, param)
        xor_last = xor_count(render(pos)[1])
    return JointFitReport(pos.detach(), xor_first, xor_last, ms)


def joint_fit_inputs(num_strands: int = 10000, num_vertices: int = 64, clusters: int = 300,
                     num_quads: int = 15000, width: float = 0.1, seed: int = 0, device="cuda"):
    """Synthetic inputs of the joint fit at 3_fitting.py's scale: seeded
    strands (cm), k-means clusters of the roots, and per cluster its mean
    strand as guide, swept into two crossed 2-row strips (2_texture.py's
    defaults) with quads shared in proportion to guide length, on the GPU
    (cardmesh).  Returns a dict of device tensors for ``joint_fit``."""
    from . import cardmesh

    dev = torch.device(device)
    strands = synth.generate_strands(num_strands, num_vertices, 0.004, seed) * 100.0
    labels = synth.kmeans_clusters(strands[:, 0], clusters, seed)
    ids = np.unique(labels)
    guides, lengths = [], []
    for c in ids:
        g = strands[labels == c].mean(0)
        guides.append(g)
        lengths.append(np.linalg.norm(np.diff(g, axis=0), axis=1).sum())
    lengths = np.array(lengths)
    quads = np.maximum(1, np.rint(num_quads / 2 * lengths / lengths.sum())).astype(int)
    guides = [synth.resample_by_arclength(g[None], int(q) + 1)[0].astype(np.float32)
              for g, q in zip(guides, quads)]
    pts, cnt, offs, _ = cardmesh._ragged(guides, dev)
    card_pos = cardmesh.strands_to_strips_device(pts, cnt, offs, 2, width, [-45.0, 45.0])
    rows = np.full(2 * len(guides), 2)
    cols = np.repeat(quads + 1, 2)
    topo = cardmesh.grid_mesh_topology_device(rows, cols, grid2cluster=np.repeat(ids, 2),
                                              device=dev)
    points = torch.from_numpy(strands.reshape(-1, 3).astype(np.float32)).to(dev)
    edges = torch.from_numpy(synth.strand_edges(num_strands, num_vertices)).to(dev)
    e2c = torch.from_numpy(np.repeat(labels, num_vertices - 1).astype(np.int32)).to(dev)
    return dict(points=points, edges=edges, edge2cluster=e2c, cluster_count=int(clusters),
                card_pos=card_pos, tri=topo["faces"],
                tri2cluster=topo["face2cluster"].to(torch.int32), quads=topo["quads"],
                h_edges=topo["h_edges"], w_edges=topo["w_edges"])
