"""The autotuner's knobs change launch shapes only: every fitting chunk and
job order gives bit-identical jobs, every DDA tile side the same mask."""
import numpy as np
import pytest
import torch

import oracle
from conftest import report
from strns2cards_amd import _core, autotune, fitting, synth

pytestmark = pytest.mark.gpu


def test_fit_chunk_and_order_do_not_change_results(dev):
    report("autotune: fit chunk x order give identical jobs")
    jobs = fitting.make_jobs(24, seed=9)
    hyper = fitting.FitHyper()
    ref = None
    for chunk in autotune.FIT_CHUNKS:
        for order in autotune.FIT_ORDERS:
            srv = fitting.CardFitServer(32, hyper, dev, chunk=chunk, order=order)
            slots = srv.admit(jobs)
            srv.run(7)
            torch.cuda.synchronize()
            got = (np.concatenate([srv.job_vertices(s) for s in slots]),
                   srv.loss.cpu().numpy()[slots])
            if ref is None:
                ref = got
            else:
                np.testing.assert_array_equal(got[0], ref[0])
                np.testing.assert_array_equal(got[1], ref[1])


def test_dda_tile_sides_give_the_same_mask(dev):
    report("autotune: every DDA tile side gives the oracle's mask")
    w = synth.small_workload()
    want = oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, 64)
    pos = torch.from_numpy(w.pos).to(dev)
    edges = torch.from_numpy(w.edges).to(dev)
    e2c = torch.from_numpy(w.edge2cluster).to(dev)
    bits = torch.empty((1, w.H, w.W, 2), dtype=torch.int32, device=dev)
    for lg in autotune.dda_tile_candidates(2):
        bits.fill_(-1)
        _core.dda_cluster_mask(pos, edges, e2c, bits, lg)
        np.testing.assert_array_equal(bits.cpu().numpy(), want)
    with pytest.raises(RuntimeError):  # 128 x 128 x 2 words does not fit LDS
        _core.dda_cluster_mask(pos, edges, e2c, bits, 7)


@pytest.mark.parametrize("K", [1024, 96, 256])
def test_dda_large_tile_bitsets_match_the_oracle(dev, K):
    """Tile sides whose bitset exceeds 32 KB resolve in the 128 KB LDS
    variant: S = 32 at side 32, S = 3 (scalar stores) and S = 8 at side 64."""
    S = (K + 31) // 32
    report(f"autotune: large tile bitsets, K={K} (S={S})")
    w = synth.small_workload()
    rng = np.random.default_rng(K)
    e2c = rng.integers(0, K, w.edges.shape[0]).astype(np.int32)
    want = oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, e2c, K)
    pos = torch.from_numpy(w.pos).to(dev)
    edges = torch.from_numpy(w.edges).to(dev)
    e2c_d = torch.from_numpy(e2c).to(dev)
    bits = torch.empty((1, w.H, w.W, S), dtype=torch.int32, device=dev)
    sides = [lg for lg in autotune.dda_tile_candidates(S) if lg >= 4]
    assert any((1 << (2 * lg)) * S * 4 > 32768 for lg in sides)
    for lg in sides:
        bits.fill_(-1)
        _core.dda_cluster_mask(pos, edges, e2c_d, bits, lg)
        np.testing.assert_array_equal(bits.cpu().numpy(), want)


def test_tuners_return_a_candidate(dev):
    report("autotune: tuners return timed candidates")
    jobs = fitting.make_jobs(16, seed=3)
    t = autotune.tune_fit(jobs, fitting.FitHyper(), dev, iters=3, reps=1)
    assert t.chunk in autotune.FIT_CHUNKS and t.order in autotune.FIT_ORDERS
    assert len(t.ms) == len(autotune.FIT_CHUNKS) * len(autotune.FIT_ORDERS)
    w = synth.small_workload()
    pos = torch.from_numpy(w.pos).to(dev)
    edges = torch.from_numpy(w.edges).to(dev)
    e2c = torch.from_numpy(w.edge2cluster).to(dev)
    bits = torch.empty((1, w.H, w.W, 2), dtype=torch.int32, device=dev)
    d = autotune.tune_dda(pos, edges, e2c, bits, reps=1)
    assert d.tile_log2 in autotune.dda_tile_candidates(2)
    assert autotune.dda_tile_for(1, w.H, w.W, 2, w.edges.shape[0]) == d.tile_log2
    # the Python API now uses the tuned side and still matches the oracle
    import strns2cards_amd as s2c
    got = s2c.dda_cluster_mask((w.H, w.W), pos, edges, e2c, 64).cpu().numpy()
    np.testing.assert_array_equal(got, oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges,
                                                               w.edge2cluster, 64))
