"""Optimizer pieces of the card-fitting loop.

Mirror of the reference module pystrns2cards/optimizer.py: ``UniformAdam``
(optimizer.py:3-68) and ``laplacian_edges`` (optimizer.py:70-99).  These run
as plain torch ops; the batched fitting kernel re-implements the same step
on the GPU for many jobs at once.
"""
import torch


class UniformAdam(torch.optim.Optimizer):
    """Adam with an isotropic second moment (one scalar per point: the sum of
    squared gradient components) and a single shared denominator, the
    largest bias-corrected RMS over all points (optimizer.py:33-68):

        g1 <- b1 g1 + (1 - b1) g
        g2 <- b2 g2 + (1 - b2) sum_c g_c^2
        p  <- p - lr * (g1 / (1 - b1^t)) / (max sqrt(g2 / (1 - b2^t)) + 1e-8)
    """

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999)):
        super().__init__(params, dict(lr=lr, betas=betas))

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
