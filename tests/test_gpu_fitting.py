"""GPU tests of batched / served card fitting against the plain per-job loop
(the drop-in API) and the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle
from conftest import report
from strns2cards_amd import fitting

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs():
    return fitting.make_jobs(12, seed=5)


def test_targets_bit_exact(dev, jobs):
    report("fit: target masks vs oracle DDA mask")
    srv = fitting.CardFitServer(16, fitting.FitHyper(), dev)
    slots = srv.admit(jobs)
    torch.cuda.synchronize()
    for s, j in zip(slots, jobs):
        n, nv = j.strands.shape[:2]
        from strns2cards_amd import synth
        want = oracle.dda_cluster_mask((j.res, j.res), j.strand_clip[None], synth.strand_edges(n, nv),
                                       np.zeros(n * (nv - 1), np.int32), 1)[0, :, :, 0] != 0
        np.testing.assert_array_equal(srv.target_mask(s), want)


def test_one_iteration_matches_plain_loop(dev, jobs):
    """After one iteration the served kernel and the plain drop-in loop agree
    (same masks, same gradient terms; only summation order differs)."""
    report("fit: one served iteration vs plain loop")
    hyper = fitting.FitHyper()
    srv = fitting.CardFitServer(16, hyper, dev)
    slots = srv.admit(jobs[:6])
    srv.run(1)
    torch.cuda.synchronize()
    for s, j in zip(slots, jobs[:6]):
        got = srv.job_vertices(s)
        want, _, _ = fitting.plain_fit(j, hyper, dev, iters=1)
        moved = np.abs(want - j.verts).max()
        assert moved > 0
        err = np.abs(got - want).max()
        assert err <= 1e-3 * moved + 1e-6, (j.jid, err, moved)


def test_alone_and_mixed_are_identical(dev, jobs):
    """A job's result does not depend on the other jobs in the batch."""
    report("fit: alone vs in a mix (bit-identical)")
    hyper = fitting.FitHyper()
    a = fitting.CardFitServer(4, hyper, dev)
    sa = a.admit([jobs[3]])[0]
    a.run(25)
    b = fitting.CardFitServer(32, hyper, dev)
    sb = b.admit(jobs[:3] + jobs[4:])  # other jobs first, so jobs[3] lands in another slot
    sb = b.admit([jobs[3]])[0]
    b.run(10)
    b.run(15)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.job_vertices(sa), b.job_vertices(sb))
    assert a.loss[sa].item() == b.loss[sb].item()


def test_final_loss_matches_plain_loop(dev, jobs):
    """100 iterations: the coverage loss (XOR pixels / target pixels) of the
    served fit is within 1% of the plain loop's, and fitting reduces it."""
    report("fit: final coverage loss served vs plain (100 iterations, 6 jobs)")
    hyper = fitting.FitHyper()
    sel = jobs[:6]
    srv = fitting.CardFitServer(8, hyper, dev)
    slots = srv.admit(sel)
    srv.run(0)
    torch.cuda.synchronize()
    init = srv.loss.cpu().numpy()[slots].sum()
    srv.run(hyper.iters, graph=True)
    torch.cuda.synchronize()
    served = srv.loss.cpu().numpy()[slots].sum()
    plain, tpix = 0, 0
    for j in sel:
        _, xor, tp = fitting.plain_fit(j, hyper, dev)
        plain += xor
        tpix += tp
    rel = abs(served - plain) / max(plain, 1)
    assert served < init
    assert rel <= 0.01, (served, plain, tpix)


def test_serve_continuous_batching(dev):
    report("fit: serve() with Poisson arrivals")
    hyper = fitting.FitHyper(iters=20)
    js = fitting.poisson_arrivals(fitting.make_jobs(24, seed=9), rate=400.0, seed=1)
    srv = fitting.CardFitServer(8, hyper, dev)
    rep = fitting.serve(js, srv, chunk=10)
    assert rep.completed == 24 and rep.iterations == 24 * 20
    assert len(rep.latencies_ms) == 24 and min(rep.latencies_ms) > 0
    # each served result equals the job run alone
    alone = fitting.CardFitServer(2, hyper, dev)
    s = alone.admit([js[5]])[0]
    alone.run(20)
    torch.cuda.synchronize()
    assert alone.loss[s].item() == rep.losses[js[5].jid]


def test_joint_fit_reduces_the_mask_error(dev):
    """3_fitting.py's loop (13 views, large steps, antialiased loss,
    planarity and Laplacian terms) on a small model lowers the XOR count
    between the card masks and the strand masks."""
    print("[joint fit] 1500 strands, 24 clusters, 64^2 x 13 views", flush=True)
    inp = fitting.joint_fit_inputs(1500, 32, 24, 1200, 0.1, seed=3, device=dev)
    rep = fitting.joint_fit(**inp, resolution=64, num_iters=30)
    assert torch.isfinite(rep.pos).all()
    assert rep.xor_last < rep.xor_first, (rep.xor_first, rep.xor_last)


def test_joint_fit_fused_follows_the_same_path(dev):
    """The fused card masks are the script's masks and the grid Laplacian is
    its sparse Laplacian, so the fit takes the same steps (to the rounding
    of the antialias gradient's atomics and of the Laplacian sums)."""
    inp = fitting.joint_fit_inputs(800, 24, 12, 600, 0.1, seed=5, device=dev)
    a = fitting.joint_fit(**inp, resolution=48, num_iters=3)
    b = fitting.joint_fit(**inp, resolution=48, num_iters=3, fused=True)
    assert a.xor_first == b.xor_first
    assert torch.allclose(a.pos, b.pos, rtol=0, atol=1e-3)

