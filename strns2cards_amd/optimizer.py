"""Optimizer pieces of the card-fitting loop.

Mirror of the reference module pystrns2cards/optimizer.py: ``UniformAdam``
(optimizer.py:3-68) and ``laplacian_edges`` (optimizer.py:70-99).  These run
as plain torch ops, or ``UniformAdam(fused=True)`` as one HIP step; the
batched fitting kernel re-implements the same step on the GPU for many jobs
at once.

``GridSystem``, ``to_differential`` and ``from_differential`` stand in for
largesteps.parameterize on the grid meshes the fitting optimizes
(3_fitting.py:347-384), with the solve in a HIP kernel.
"""
import numpy as np
import torch

from . import _core


class UniformAdam(torch.optim.Optimizer):
    """Adam with an isotropic second moment (one scalar per point: the sum of
    squared gradient components) and a single shared denominator, the
    largest bias-corrected RMS over all points (optimizer.py:33-68):

        g1 <- b1 g1 + (1 - b1) g
        g2 <- b2 g2 + (1 - b2) sum_c g_c^2
        p  <- p - lr * (g1 / (1 - b1^t)) / (max sqrt(g2 / (1 - b2^t)) + 1e-8)

    ``fused=True`` (no reference counterpart) takes each step on the GPU in
    two launches (``s2c_uniform_adam_step``: moments with per-workgroup
    maxima, then the shared denominator and the update) instead of a dozen
    tensor operations; parameters must then be float32 contiguous GPU
    tensors.  The state tensors are the same.
    """

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), fused: bool = False):
        super().__init__(params, dict(lr=lr, betas=betas))
        self.fused = fused

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                g = p.grad
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["g1"] = torch.zeros_like(p)
                    st["g2"] = torch.zeros_like(p[..., :1])
                st["step"] += 1
                t = st["step"]
                if self.fused:
                    if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous()
                            and g.is_contiguous() and p.dim() >= 1):
                        raise ValueError("UniformAdam(fused=True) needs float32 contiguous "
                                         "GPU parameters and gradients")
                    dim = p.shape[-1]  # g2 holds one entry per row of the last axis
                    _core.uniform_adam_step(p.view(-1, dim), g.view(-1, dim),
                                            st["g1"].view(-1, dim), st["g2"].view(-1),
                                            float(lr), float(b1), float(b2), int(t))
                    continue
                st["g1"].mul_(b1).add_(g, alpha=1 - b1)
                st["g2"].mul_(b2).add_(g.square().sum(dim=-1, keepdim=True), alpha=1 - b2)
                m1 = st["g1"] / (1 - b1 ** t)
                m2 = st["g2"] / (1 - b2 ** t)
                denom = m2.sqrt().max() + 1e-8
                p.sub_(m1 / denom, alpha=lr)


def laplacian_edges(verts: torch.Tensor, edges: torch.Tensor) -> torch.Tensor:
    """Uniform graph Laplacian (V, V) as a sparse COO tensor: -1 per distinct
    neighbour pair, degree on the diagonal (optimizer.py:70-99)."""
    V = verts.shape[0]
    i = edges[:, 0].long()
    j = edges[:, 1].long()
    adj = torch.stack([torch.cat([i, j]), torch.cat([j, i])], 0).unique(dim=1)
    ones = torch.ones(adj.shape[1], device=verts.device, dtype=torch.float32)
    idx = torch.cat([adj, torch.stack([adj[0], adj[0]], 0)], 1)
    vals = torch.cat([-ones, ones])
    return torch.sparse_coo_tensor(idx, vals, (V, V)).coalesce()


# ------------------------------------------------------------------------
# Large-steps parameterization of grid meshes
# ------------------------------------------------------------------------

class GridSystem:
    """The system matrix M = I + lam * (L_h + w_weight * L_w) of the card
    fitting (3_fitting.py:347-356) for concatenated grid meshes
    (texture_utils.py:525-655): grid g has ``rows[g]`` rows of ``cols[g]``
    vertices, vertex (i, j) at ``voff[g] + i * cols[g] + j``, h_edges join
    (i, j)-(i, j+1) and w_edges (i, j)-(i+1, j); L_h, L_w are their uniform
    Laplacians (``laplacian_edges``).

    It stands where the reference passes the sparse M to
    largesteps.parameterize: ``to_differential(M, x)`` is u = M x and
    ``from_differential(M, u)`` solves M x = u on the GPU.  As the
    reference factors M once (cholespy) and solves with the factor every
    iteration, the constructor computes the factor on the GPU (the pivots of
    each grid shape's tridiagonal systems after a DCT across its rows) and
    every solve reuses it.  On a non-GPU device only the layout is built
    (host logic); apply and solve need the GPU.
    """

    def __init__(self, rows, cols, lam: float = 100.0, w_weight: float = 1.0 / 100,
                 device="cuda"):
        rows = np.asarray(rows, dtype=np.int64).reshape(-1)
        cols = np.asarray(cols, dtype=np.int64).reshape(-1)
        if rows.shape != cols.shape:
            raise ValueError("GridSystem: rows and cols must have one entry per grid")
        if rows.size and (rows.min() < 1 or cols.min() < 1):
            raise ValueError("GridSystem: every grid needs at least one row and one column")
        if lam < 0 or w_weight < 0:
            raise ValueError("GridSystem: lam and w_weight must be >= 0")
        self.lam, self.w_weight = float(lam), float(w_weight)
        verts = rows * cols
        voff = np.zeros_like(verts)
        if verts.size:
            voff[1:] = np.cumsum(verts)[:-1]
        self.num_verts = int(verts.sum())
        if self.num_verts >= 2**31:
            raise ValueError("GridSystem: more than 2^31 - 1 vertices")
        self.rows_host, self.cols_host = rows.astype(np.int32), cols.astype(np.int32)
        self.voff_host = voff.astype(np.int32)
        self.max_verts = int(verts.max()) if verts.size else 1
        self.max_rows = int(rows.max()) if rows.size else 1
        dev = torch.device(device)
        self.rows = torch.from_numpy(self.rows_host).to(dev)
        self.cols = torch.from_numpy(self.cols_host).to(dev)
        self.voff = torch.from_numpy(self.voff_host).to(dev)
        # the factor (pivots of each grid's decoupled tridiagonal systems)
        # depends on the grid shape only: one table per distinct (rows, cols)
        shapes, inverse = np.unique(np.stack([rows, cols], 1).reshape(-1, 2), axis=0,
                                    return_inverse=True)
        sizes = shapes[:, 0] * shapes[:, 1]
        starts = np.zeros(len(shapes), np.int64)
        if len(shapes):
            starts[1:] = np.cumsum(sizes)[:-1]
        self.fac_off_host = starts[inverse.reshape(-1)].astype(np.int32)
        self.fac_off = torch.from_numpy(self.fac_off_host).to(dev)
        self.factor = torch.empty(int(sizes.sum()), dtype=torch.float32, device=dev)
        if len(shapes) and dev.type == "cuda":
            shape_rows = torch.from_numpy(shapes[:, 0].astype(np.int32)).to(dev)
            shape_cols = torch.from_numpy(shapes[:, 1].astype(np.int32)).to(dev)
            shape_off = torch.from_numpy(starts.astype(np.int32)).to(dev)
            _core.grid_system_factor(shape_rows, shape_cols, shape_off, self.lam, self.w_weight,
                                     self.factor)

    @property
    def num_grids(self) -> int:
        return int(self.rows_host.shape[0])

    @property
    def shape(self):
        return (self.num_verts, self.num_verts)

    @classmethod
    def from_edges(cls, num_verts: int, h_edges, w_edges, lam: float = 100.0,
                   w_weight: float = 1.0 / 100, device="cuda") -> "GridSystem":
        """Recovers the grid layout from the h_edges / w_edges of
        combine_wisps_with_cluster_labels (texture_utils.py:597-655) and
        checks that the edges are exactly those of that layout (ValueError
        otherwise: M would not be a grid system)."""
        V = int(num_verts)
        h = _undirected_keys(h_edges, V, "h_edges")
        w = _undirected_keys(w_edges, V, "w_edges")
        if V == 0:
            if h.size or w.size:
                raise ValueError("GridSystem.from_edges: edges given for zero vertices")
            return cls([], [], lam, w_weight, device)
        has_h = np.zeros(V, dtype=bool)  # edge v -> v + 1 inside a row
        hh = h[(h % V) == (h // V) + 1]
        has_h[hh // V] = True
        starts = np.flatnonzero(np.concatenate([[True], ~has_h[:-1]]))
        lens = np.diff(np.concatenate([starts, [V]]))
        # row r continues the grid of row r - 1 when it is as long and the
        # first vertices are joined by a w_edge
        linked = np.zeros(starts.size, dtype=bool)
        if starts.size > 1:
            same = lens[1:] == lens[:-1]
            key = starts[:-1] * V + starts[1:]
            linked[1:] = same & np.isin(key, w)
        first = np.flatnonzero(~linked)
        rows = np.diff(np.concatenate([first, [starts.size]]))
        cols = lens[first]
        sys_ = cls(rows, cols, lam, w_weight, device)
        eh, ew = grid_edges(sys_.rows_host, sys_.cols_host)
        if not (np.array_equal(_undirected_keys(eh, V, "h"), h)
                and np.array_equal(_undirected_keys(ew, V, "w"), w)):
            raise ValueError("GridSystem.from_edges: the edges are not those of concatenated "
                             "grid meshes (texture_utils.wisp_to_grid_mesh)")
        return sys_

    def quad_layout(self):
        """(per-grid quad offsets (device int32), total quads) of the grids'
        (rows - 1)(cols - 1) quads, concatenated."""
        if not hasattr(self, "_qoff"):
            nq = (self.rows_host.astype(np.int64) - 1) * (self.cols_host.astype(np.int64) - 1)
            off = np.zeros_like(nq)
            if nq.size:
                off[1:] = np.cumsum(nq)[:-1]
            self._qoff = torch.from_numpy(off.astype(np.int32)).to(self.rows.device)
            self._nq = int(nq.sum())
        return self._qoff, self._nq

    def _check(self, t: torch.Tensor, who: str) -> torch.Tensor:
        if t.dim() != 2 or t.shape[0] != self.num_verts or t.shape[1] != 3:
            raise ValueError(f"{who}: expected a ({self.num_verts}, 3) tensor, got "
                             f"{tuple(t.shape)}")
        if t.device != self.rows.device:
            raise ValueError(f"{who}: tensor on {t.device}, system on {self.rows.device}")
        return t.contiguous().float()

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """u = M x (no autograd)."""
        x = self._check(x, "GridSystem.apply")
        u = torch.empty_like(x)
        _core.grid_system_apply(self.rows, self.cols, self.voff, self.lam, self.w_weight, x, u)
        return u

    def solve(self, u: torch.Tensor) -> torch.Tensor:
        """x = M^{-1} u (no autograd)."""
        u = self._check(u, "GridSystem.solve")
        x = torch.empty_like(u)
        _core.grid_system_solve(self.rows, self.cols, self.voff, self.fac_off, self.factor,
                                self.max_verts, self.max_rows, self.lam, u, x)
        return x


def grid_laplacian(M: GridSystem, x: torch.Tensor, w_weight: float = 1.0) -> torch.Tensor:
    """(L_h + w_weight L_w) x on M's grids, differentiable (L is symmetric,
    so its backward is itself).  With w_weight = 1 it is
    laplacian_edges(x, cat(h_edges, w_edges)) @ x, the uniform Laplacian of
    3_fitting.py:359, without a sparse matrix."""
    if not isinstance(M, GridSystem):
        raise TypeError("grid_laplacian: M must be a GridSystem")
    return _GridLaplacian.apply(x, M, float(w_weight))


class _GridLaplacian(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, M, ww):
        ctx.M, ctx.ww = M, ww
        x = M._check(x, "grid_laplacian")
        y = torch.empty_like(x)
        _core.grid_laplacian(M.rows, M.cols, M.voff, ww, x, y)
        return y

    @staticmethod
    def backward(ctx, grad):
        g = ctx.M._check(grad, "grid_laplacian")
        y = torch.empty_like(g)
        _core.grid_laplacian(ctx.M.rows, ctx.M.cols, ctx.M.voff, ctx.ww, g, y)
        return y, None, None


def grid_quad_diagonals(M: GridSystem, pos: torch.Tensor) -> torch.Tensor:
    """(Q, 3) (p0 + p2) - (p1 + p3) of every quad of M's grids, in the quad
    order of wisp_to_grid_mesh (texture_utils.py:557-565), differentiable:
    sum(D ** 2) is the planarity loss of 3_fitting.py:181-199.  The backward
    gathers each vertex's at most four quads in a fixed order (no
    accumulating scatter, no index sort)."""
    if not isinstance(M, GridSystem):
        raise TypeError("grid_quad_diagonals: M must be a GridSystem")
    return _GridQuadDiagonals.apply(pos, M)


class _GridQuadDiagonals(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, M):
        ctx.M = M
        pos = M._check(pos, "grid_quad_diagonals")
        qoff, nq = M.quad_layout()
        D = torch.empty((nq, 3), dtype=torch.float32, device=pos.device)
        _core.grid_quad_diagonals(M.rows, M.cols, M.voff, qoff, pos, D, False)
        return D

    @staticmethod
    def backward(ctx, grad):
        M = ctx.M
        qoff, _ = M.quad_layout()
        dpos = torch.empty((M.num_verts, 3), dtype=torch.float32, device=grad.device)
        _core.grid_quad_diagonals(M.rows, M.cols, M.voff, qoff, dpos, grad.contiguous().float(),
                                  True)
        return dpos, None


def grid_quads(rows, cols) -> np.ndarray:
    """(Q, 4) quads [i0, i1, i2, i3] of concatenated grid meshes, in the order
    of wisp_to_grid_mesh (texture_utils.py:557-565) with the vertex offsets
    of combine_wisps_with_cluster_labels (texture_utils.py:634-655): the
    quads grid_quad_diagonals differentiates."""
    qs = []
    off = 0
    for S, N in zip(np.asarray(rows).tolist(), np.asarray(cols).tolist()):
        idx = off + np.arange(S * N, dtype=np.int64).reshape(S, N)
        qs.append(np.stack([idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]], -1).reshape(-1, 4))
        off += S * N
    return np.concatenate(qs, 0) if qs else np.zeros((0, 4), np.int64)


def grid_edges(rows, cols):
    """h_edges and w_edges of concatenated grid meshes, in the order of
    wisp_to_grid_mesh (texture_utils.py:566-584) with the vertex offsets of
    combine_wisps_with_cluster_labels (texture_utils.py:634-655)."""
    hs, ws = [], []
    off = 0
    for S, N in zip(np.asarray(rows).tolist(), np.asarray(cols).tolist()):
        idx = off + np.arange(S * N, dtype=np.int64).reshape(S, N)
        hs.append(np.stack([idx[:, :-1].reshape(-1), idx[:, 1:].reshape(-1)], 1))
        ws.append(np.stack([idx[:-1, :].reshape(-1), idx[1:, :].reshape(-1)], 1))
        off += S * N
    empty = np.zeros((0, 2), dtype=np.int64)
    return (np.concatenate(hs) if hs else empty), (np.concatenate(ws) if ws else empty)


def _undirected_keys(edges, V: int, who: str) -> np.ndarray:
    e = edges.detach().cpu().numpy() if isinstance(edges, torch.Tensor) else np.asarray(edges)
    e = e.astype(np.int64).reshape(-1, 2)
    if e.size and (e.min() < 0 or e.max() >= V):
        raise ValueError(f"GridSystem.from_edges: {who} index out of range")
    e = e[e[:, 0] != e[:, 1]]  # self-loops add nothing to a graph Laplacian
    lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
    return np.unique(lo * V + hi)


class _FromDifferential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, M):
        ctx.M = M
        return M.solve(u)

    @staticmethod
    def backward(ctx, grad):
        return ctx.M.solve(grad), None  # M is symmetric


class _ToDifferential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, M):
        ctx.M = M
        return M.apply(x)

    @staticmethod
    def backward(ctx, grad):
        return ctx.M.apply(grad), None


def to_differential(M: GridSystem, v: torch.Tensor) -> torch.Tensor:
    """u = M v (largesteps.parameterize.to_differential, called at
    3_fitting.py:365)."""
    if not isinstance(M, GridSystem):
        raise TypeError("to_differential: M must be a GridSystem")
    return _ToDifferential.apply(v, M)


def from_differential(M: GridSystem, u: torch.Tensor, method: str = "Cholesky") -> torch.Tensor:
    """v = M^{-1} u, differentiable (largesteps.parameterize.from_differential,
    called at 3_fitting.py:384 and 457).  `method` is accepted for the
    reference's signature ('Cholesky' or 'CG'); the grid solve is direct
    either way."""
    if not isinstance(M, GridSystem):
        raise TypeError("from_differential: M must be a GridSystem")
    if method not in ("Cholesky", "CG"):
        raise ValueError(f"from_differential: unknown method {method!r}")
    return _FromDifferential.apply(u, M)
