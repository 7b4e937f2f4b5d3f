"""Autotuner for the launch shapes of the batched fitting kernel and the
binned DDA -> cluster-mask pass.

Every knob tuned here changes only how the work is laid out on the GPU,
never a result (each job's sums are in a fixed order; OR is order-free),
so the tuner simply times the candidates on the device it runs on and keeps
the fastest:

* fitting (``tune_fit``): ``chunk``, the silhouette samples a workgroup
  evaluates per pass (256 / 512 / 1024: LDS per workgroup, hence occupancy,
  against passes per iteration), and the launch ``order`` of the jobs
  ("admission", or "cost": most work first, so the long jobs of a mixed
  batch do not start last);
* DDA -> mask (``tune_dda``): the screen-tile side of the binning (every
  power of two whose tile bitset fits 128 KB of LDS), which trades
  (workgroup, tile) runs and their atomics in the stage pass against tile
  workgroups and their balance in the resolve.

Results are kept per shape in this process; ``dda_cluster_mask`` uses the
tuned tile side of its shape when one exists.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import _core

FIT_CHUNKS = (256, 512, 1024)
FIT_ORDERS = ("admission", "cost")

_DDA_TILES: dict[tuple, int] = {}


def _time(fn, reps: int) -> float:
    """Median milliseconds of fn() over reps runs (HIP events on the current
    stream, after one untimed run)."""
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


# ------------------------------------------------------------------ fitting


@dataclass
class FitTuning:
    chunk: int
    order: str
    ms: dict = field(default_factory=dict)  # "chunk/order" -> launch ms


def tune_fit(jobs, hyper, dev, iters: int = 10, reps: int = 3) -> FitTuning:
    """Times one ``iters``-iteration launch over ``jobs`` (a representative
    sample of the workload) for every (chunk, order) and returns the
    fastest.  Each candidate starts from a freshly admitted table, so all
    see the same state."""
    from .fitting import CardFitServer

    ms = {}
    for chunk in FIT_CHUNKS:
        for order in FIT_ORDERS:
            srv = CardFitServer(len(jobs), hyper, dev, chunk=chunk, order=order)
            srv.admit(jobs)
            ms[f"{chunk}/{order}"] = _time(lambda: srv.run(iters), reps)
            del srv
    best = min(ms, key=ms.get)
    chunk, order = best.split("/")
    return FitTuning(int(chunk), order, ms)


# ------------------------------------------------------------------ DDA -> mask


def dda_tile_candidates(S: int) -> list[int]:
    """Tile sides (log2) whose (side^2 x S) word bitset fits 128 KB of LDS
    (sides above the default, whose bitset fits 32 KB, run one resolve
    workgroup per CU)."""
    return [lg for lg in range(2, 7) if (1 << (2 * lg)) * S * 4 <= 131072]


def _dda_key(B, H, W, S, E):
    return (int(B), int(H), int(W), int(S), int(E).bit_length())


def dda_tile_for(B: int, H: int, W: int, S: int, E: int) -> int:
    """Tuned tile side (log2) for this shape, or 0 (automatic)."""
    return _DDA_TILES.get(_dda_key(B, H, W, S, E), 0)


@dataclass
class DdaTuning:
    tile_log2: int
    ms: dict = field(default_factory=dict)  # tile side -> ms per call


def tune_dda(pos: torch.Tensor, edges: torch.Tensor, edge2cluster: torch.Tensor,
             bitset: torch.Tensor, reps: int = 5) -> DdaTuning:
    """Times ``dda_cluster_mask`` into ``bitset`` (B, H, W, S) for every
    tile side that fits LDS, records the fastest for this shape and
    returns it with the timings."""
    B, H, W, S = bitset.shape
    ms = {}
    for lg in dda_tile_candidates(S):
        ms[1 << lg] = _time(lambda: _core.dda_cluster_mask(pos, edges, edge2cluster, bitset, lg),
                            reps)
    side = min(ms, key=ms.get)
    lg = side.bit_length() - 1
    _DDA_TILES[_dda_key(B, H, W, S, edges.shape[0])] = lg
    return DdaTuning(lg, ms)
