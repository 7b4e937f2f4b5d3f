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


def cpu_job(j):
    from oracle import cardfit

    n, nv = j.strands.shape[:2]
    return cardfit.job_dict(j.res, j.quads, j.strand_clip.reshape(n, nv, 4), j.mvp, j.verts)


def test_iterations_match_the_cpu_loop(dev, jobs):
    """The batched kernel against the oracle's fitting loop (3_fitting.py:
    381-407 restated on the CPU in float64 with a Cholesky solve of the
    largesteps system): after 1 and after 3 iterations every vertex agrees
    to 1e-3 of how far the job has moved plus the float32 rounding of the
    solve, cond(M) * 2^-24 * max|vertex| (positions are recovered from
    parameters M x whose magnitude is |M| |x|; the masks are the same)."""
    from oracle import cardfit

    report("fit: 1 and 3 iterations vs the CPU fitting loop")
    hyper = fitting.FitHyper()
    hyp = cardfit.hyper_dict(hyper)
    for iters in (1, 3):
        srv = fitting.CardFitServer(16, hyper, dev)
        slots = srv.admit(jobs[:8])
        srv.run(iters)
        torch.cuda.synchronize()
        for s, j in zip(slots, jobs[:8]):
            job = cpu_job(j)
            want, _, _ = cardfit.fit_job(job, hyp, iters)
            got = srv.job_vertices(s)
            M = oracle.largesteps_matrix(want.shape[0], job["h_edges"], job["w_edges"], hyper.lam,
                                         hyper.w_weight)
            solve_eps = np.linalg.cond(M) * 2.0 ** -24 * np.abs(j.verts).max()
            moved = np.abs(want - j.verts).max()
            err = np.abs(got - want).max()
            report(f"  job {j.jid} res {j.res} Q {j.quads} iters {iters}: moved {moved:.3e} "
                   f"err {err:.3e} (solve rounding {solve_eps:.2e})")
            assert moved > 10 * solve_eps
            assert err <= 1e-3 * moved + solve_eps, (j.jid, iters, err, moved, solve_eps)


def test_final_loss_matches_the_cpu_loop(dev):
    """All 100 iterations of 30 jobs (two of every resolution x quads class):
    the summed final coverage loss of the batched kernel is within 1% of the
    oracle's CPU fitting loop (the north-star's 'equal final loss')."""
    from oracle import cardfit

    report("fit: 100 iterations, 30 stratified jobs vs the CPU fitting loop")
    hyper = fitting.FitHyper()
    hyp = cardfit.hyper_dict(hyper)
    sample = stratified(fitting.make_jobs(400, seed=11), 2)
    assert len(sample) == 30
    srv = fitting.CardFitServer(32, hyper, dev)
    slots = srv.admit(sample)
    srv.run(hyper.iters)
    torch.cuda.synchronize()
    loss = srv.loss.cpu().numpy()
    gpu = cpu = tpix = 0
    for s, j in zip(slots, sample):
        _, xor, tp = cardfit.fit_job(cpu_job(j), hyp, hyper.iters)
        gpu += float(loss[s])
        cpu += xor
        tpix += tp
    rel = abs(gpu - cpu) / max(cpu, 1)
    report(f"  coverage loss GPU {gpu / tpix:.5f} CPU {cpu / tpix:.5f} (rel diff {rel:.4f})")
    assert rel <= 0.01, (gpu, cpu, tpix)


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


def test_measuring_changes_nothing(dev, jobs):
    """run(0) (also the warm-up before a graph capture) writes only the loss:
    a job measured first, then run in graph-replayed chunks, ends
    bit-identical to the same job run in one launch."""
    report("fit: run(0) and graph replays leave the trajectory unchanged")
    hyper = fitting.FitHyper(iters=40)
    a = fitting.CardFitServer(2, hyper, dev)
    sa = a.admit([jobs[2]])[0]
    a.run(40)
    b = fitting.CardFitServer(2, hyper, dev)
    sb = b.admit([jobs[2]])[0]
    b.run(0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.job_vertices(sb), jobs[2].verts)
    for _ in range(4):
        b.run(10, graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.job_vertices(sa), b.job_vertices(sb))
    assert a.loss[sa].item() == b.loss[sb].item()


def test_iteration_limit(dev, jobs):
    """A job admitted with a limit of 30 iterations: two 20-iteration
    launches run 20 + 10 of them (the same result as one 30-iteration
    launch), and a later launch changes neither its vertices nor its loss;
    a job admitted beside it without a limit keeps advancing."""
    report("fit: per-job iteration limit")
    hyper = fitting.FitHyper()
    a = fitting.CardFitServer(4, hyper, dev)
    sa = a.admit([jobs[1]])[0]
    a.run(30)
    b = fitting.CardFitServer(4, hyper, dev)
    sb = b.admit([jobs[1]], limit=30)[0]
    sc = b.admit([jobs[2]])[0]
    b.run(20)
    b.run(20)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.job_vertices(sa), b.job_vertices(sb))
    assert a.loss[sa].item() == b.loss[sb].item()
    assert b.steps[sb].item() == 30 and b.steps[sc].item() == 40
    v, l, vc = b.job_vertices(sb), b.loss[sb].item(), b.job_vertices(sc)
    b.run(10, graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.job_vertices(sb), v)
    assert b.loss[sb].item() == l and b.steps[sb].item() == 30
    assert b.steps[sc].item() == 50 and not np.array_equal(b.job_vertices(sc), vc)


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


def run_alone(job, hyper, dev):
    """The job by itself: (final vertices, final loss)."""
    srv = fitting.CardFitServer(1, hyper, dev)
    s = srv.admit([job])[0]
    srv.run(hyper.iters)
    torch.cuda.synchronize()
    return srv.job_vertices(s), srv.loss[s].item()


def stratified(jobs, per_class):
    """Up to `per_class` jobs of every (resolution, quads) class."""
    seen, out = {}, []
    for j in jobs:
        k = (j.res, j.quads)
        if seen.get(k, 0) < per_class:
            seen[k] = seen.get(k, 0) + 1
            out.append(j)
    return out


@pytest.mark.parametrize("lam", [200.0, 1000.0, 5000.0])
def test_serve_poisson_rates(dev, lam):
    """Config 4: Poisson arrivals at lambda jobs/s, 100 iterations per job,
    continuous batching with graph replays and no per-chunk host wait;
    every completed job's vertices and loss equal the job run alone."""
    report(f"fit: serve() at lambda={lam:.0f}/s, 100 iterations per job")
    hyper = fitting.FitHyper(iters=100)
    js = fitting.poisson_arrivals(fitting.make_jobs(160, seed=int(lam) + 3), rate=lam, seed=2)
    srv = fitting.CardFitServer(256, hyper, dev)
    rep = fitting.serve(js, srv, chunk=10, keep_vertices=True)
    assert rep.completed == len(js) and rep.iterations == len(js) * 100
    assert len(rep.latencies_ms) == len(js) and min(rep.latencies_ms) > 0
    for j in js:
        v, loss = run_alone(j, hyper, dev)
        np.testing.assert_array_equal(rep.vertices[j.jid], v)
        assert rep.losses[j.jid] == loss, j.jid
    report(f"  p50 {rep.percentile(50):.2f} ms, p99 {rep.percentile(99):.2f} ms, "
           f"{rep.iters_per_s:.0f} it/s")


def test_serve_waits_on_the_budget(dev):
    """A budget with room for three 256^2 masks: arriving jobs wait for the
    pool, every job completes, and results equal the jobs run alone."""
    report("fit: serve() under a tight HBM budget")
    hyper = fitting.FitHyper(iters=20)
    js = fitting.make_jobs(40, seed=11)
    for j in js:
        j.arrival = 0.0  # all at once: admission is bounded by the pool alone
    cap = 32
    table = cap * (40 * fitting.MAX_VERTS + fitting.SLOT_SCALAR_BYTES)
    srv = fitting.CardFitServer(cap, hyper, dev, budget_bytes=table + 3 * 4 * fitting.MAX_WORDS)
    first = [j for j in js if srv.can_admit(j, [])]
    assert len(first) == len(js)  # each fits on its own ...
    pw, taken = [], 0
    for j in js:  # ... but not all at once
        if srv.can_admit(j, pw):
            pw.append(j.words)
            taken += 1
    assert taken < len(js)
    rep = fitting.serve(js, srv, chunk=10, keep_vertices=True)
    assert rep.completed == len(js)
    assert srv.pool.top <= srv.pool.words and not srv.state and srv.used_bytes == 0
    for j in js[:8]:
        v, loss = run_alone(j, hyper, dev)
        np.testing.assert_array_equal(rep.vertices[j.jid], v)
        assert rep.losses[j.jid] == loss


@pytest.mark.slow
def test_batched_4096_jobs(dev):
    """Config 3: 4,096 jobs in one persistent launch of 100 iterations.
    Every target mask is bit-exact with the oracle's DDA mask; 32 jobs
    spread over every (resolution, quads) class end bit-identical to the
    same job run alone; and on 24 such jobs the coverage loss (XOR pixels /
    target pixels, summed) is within 1% of the plain per-job loop's."""
    from strns2cards_amd import synth

    report("fit: 4,096 jobs x 100 iterations in one launch (config 3)")
    hyper = fitting.FitHyper()
    jobs = fitting.make_jobs(4096, seed=0)
    srv = fitting.CardFitServer(4096, hyper, dev)
    slots = srv.admit(jobs)
    srv.run(0)
    torch.cuda.synchronize()
    init = srv.loss.cpu().numpy()[slots]
    for s, j in zip(slots, jobs):
        n, nv = j.strands.shape[:2]
        want = oracle.dda_cluster_mask((j.res, j.res), j.strand_clip[None],
                                       synth.strand_edges(n, nv),
                                       np.zeros(n * (nv - 1), np.int32), 1)[0, :, :, 0] != 0
        np.testing.assert_array_equal(srv.target_mask(s), want, err_msg=f"job {j.jid}")
    srv.run(hyper.iters)
    torch.cuda.synchronize()
    final = srv.loss.cpu().numpy()[slots]
    assert final.sum() < init.sum()
    strat = stratified(jobs, 3)
    assert len({(j.res, j.quads) for j in strat}) == 15
    for j in strat[:32]:
        v, loss = run_alone(j, hyper, dev)
        np.testing.assert_array_equal(srv.job_vertices(slots[j.jid]), v, err_msg=f"job {j.jid}")
        assert final[j.jid] == loss
    report("  plain per-job loop on 24 stratified jobs")
    plain, served, tpix = 0, 0.0, 0
    for j in stratified(jobs, 2)[:24]:
        _, xor, tp = fitting.plain_fit(j, hyper, dev)
        plain += xor
        tpix += tp
        served += float(final[j.jid])
    rel = abs(served - plain) / max(plain, 1)
    report(f"  coverage loss served {served / tpix:.5f} plain {plain / tpix:.5f} (rel diff {rel:.4f})")
    assert rel <= 0.01, (served, plain, tpix)


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



def test_joint_fit_fused_with_reordered_quads(dev):
    """joint_fit(fused=True) given quads in another order than the grids'
    falls back to planarity_loss(pos, quads): the same loss, so the same
    path up to the rounding of the sums."""
    inp = fitting.joint_fit_inputs(800, 24, 12, 600, 0.1, seed=5, device=dev)
    a = fitting.joint_fit(**inp, resolution=48, num_iters=3, fused=True)
    perm = torch.randperm(inp["quads"].shape[0], generator=torch.Generator().manual_seed(1))
    inp["quads"] = inp["quads"][perm.to(inp["quads"].device)].contiguous()
    b = fitting.joint_fit(**inp, resolution=48, num_iters=3, fused=True)
    assert abs(a.xor_last - b.xor_last) <= 2
    assert torch.allclose(a.pos, b.pos, atol=1e-4)
