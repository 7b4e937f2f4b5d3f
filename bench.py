#!/usr/bin/env python3
"""Benchmark of the strns2cards hot path on one MI355X.

Workload (synthetic, seeded; no dataset is available): 50,000 strands x 100
vertices fp32 (step 0.003, seed 1), 1,024 k-means clusters, 4 orthographic
views at 1024 x 1024, 1,024-bit cluster bitset per pixel (32 words), plus
1,024 cards x 20 quads for the triangle / antialias stages.

Timed step: the strand DDA -> cluster-bitset pass over all 4 views
(``_core.dda_cluster_mask``: one staged walk over the edges, run ordering,
per-tile resolve), inputs resident in HBM.  ``value`` is DDA GB/s with algorithmic bytes
B*V*16 + E*8 + E*4 + B*H*W*S*4 (vertex reads per view, edges and
edge->cluster once, bitset written once) per step.  The card raster -> mask
pass and the antialias edge gradient at the same size are timed in the same
run and reported as extra fields.  ``cpu_baseline`` times the CPU oracle
(a C restatement of the reference kernels, OpenMP) on the same step, all 4
views, at 1 thread and at the host's cores (capped at 16); ``card_fit``
carries its own CPU baseline (the oracle's per-job fitting loop).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "card-fit iters/s served vs plain loop at equal loss; p50/p99 latency; DDA GB/s"


def dda_mask_bytes(B, V, E, H, W, S):
    return B * V * 16 + E * 8 + E * 4 + B * H * W * S * 4


def traffic_from_profiles():
    """HBM bytes per launch of the DDA->mask entry point, from the committed
    rocprofv3 PMC summary (None when absent)."""
    p = os.path.join(ROOT, "profiles", "dda_mask_full_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


# Worker processes and OpenMP threads for the CPU legs: the host cores this
# process may run on, capped at the 16 cores a one-GPU box gives a job.
CPU_SHARE = 16


def host_cores():
    visible = len(os.sched_getaffinity(0))
    return visible, max(1, min(CPU_SHARE, visible))


def cpu_baseline(w, seconds):
    """Oracle DDA->mask over all 4 views of the workload: one timed step at 1
    thread, and steps repeated for ~`seconds` on the host's cores."""
    import numpy as np

    import oracle

    oracle.lib()
    visible, threads = host_cores()
    S = (w.num_clusters + 31) // 32
    B, V = w.pos.shape[:2]
    nbytes = dda_mask_bytes(B, V, w.edges.shape[0], w.H, w.W, S)
    t0 = time.perf_counter()
    oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, w.num_clusters, threads=1)
    dt1 = time.perf_counter() - t0
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.dda_cluster_mask((w.H, w.W), w.pos, w.edges, w.edge2cluster, w.num_clusters,
                                threads=threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dtn = (time.perf_counter() - t0) / reps
    return {"value": round(nbytes / dtn / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port", "host_cores_visible": visible,
            "value_1_thread": round(nbytes / dt1 / 1e9, 3),
            "sample": f"the whole step (4 views, 50k strands, 1024^2, S=32): 1 rep at 1 thread "
                      f"({dt1 * 1e3:.0f} ms), {reps} rep(s) at {threads} threads "
                      f"({dtn * 1e3:.1f} ms/rep; OpenMP over views x edge ranges)"}


def fit_job_dict(job):
    """Host arrays of one fitting job for the CPU fitting loop."""
    from oracle import cardfit

    n, nv = job.strands.shape[:2]
    return cardfit.job_dict(job.res, job.quads, job.strand_clip.reshape(n, nv, 4), job.mvp,
                            job.verts)


def fit_cpu_baseline(jobs, hyper, one_thread_jobs=320, pool_jobs=2048):
    """Card fitting on the host: the oracle's per-job loop (oracle/cardfit.py,
    the iteration of 3_fitting.py:381-407 on the C oracle and numpy), all
    100 iterations per job, first in this process (1 thread) on
    `one_thread_jobs` jobs, then one single-threaded process per core on
    `pool_jobs` jobs (spawned children that only import the oracle)."""
    import multiprocessing as mp

    from oracle import cardfit

    hyp = cardfit.hyper_dict(hyper)
    visible, procs = host_cores()
    it1, s1, _, _ = cardfit.fit_jobs(([fit_job_dict(j) for j in jobs[:one_thread_jobs]], hyp,
                                      hyper.iters))
    sel = [fit_job_dict(j) for j in jobs[:pool_jobs]]
    parts = [(sel[i:i + 8], hyp, hyper.iters) for i in range(0, len(sel), 8)]  # balanced
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS")}
    os.environ["OMP_NUM_THREADS"] = os.environ["OPENBLAS_NUM_THREADS"] = "1"
    try:
        with mp.get_context("spawn").Pool(procs) as pool:
            pool.map(cardfit.fit_jobs, [([], hyp, 1)] * procs)  # start-up outside the timing
            t0 = time.perf_counter()
            res = pool.map(cardfit.fit_jobs, parts)
            wall = time.perf_counter() - t0
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    its = sum(r[0] for r in res)
    return {"value": round(its / wall, 1), "unit": "card-fit iterations/s", "cores": procs,
            "kind": "port", "host_cores_visible": visible,
            "value_1_thread": round(it1 / s1, 1),
            "sample": f"jobs 0..{one_thread_jobs - 1} x {hyper.iters} iterations in one process "
                      f"({s1:.1f} s); jobs 0..{pool_jobs - 1} x {hyper.iters} iterations over "
                      f"{procs} single-threaded processes ({wall:.1f} s wall), target masks "
                      f"included, as in the plain loop"}


def stratified(jobs, per_class):
    """Up to `per_class` jobs of every (resolution, quads) class, in order."""
    seen, out = {}, []
    for j in jobs:
        k = (j.res, j.quads)
        if seen.get(k, 0) < per_class:
            seen[k] = seen.get(k, 0) + 1
            out.append(j)
    return out


def fit_roofline_profile():
    """VALU instructions of one fit_cards_kernel launch (4,096 jobs x 100
    iterations) from the committed rocprofv3 SQ counter summary (None when
    absent)."""
    p = os.path.join(ROOT, "profiles", "fit_cards_sq_counters.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


# Wave64 VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# VALU instruction per SIMD every 4 cycles.
VALU_PEAK_GWIS = 256 * 4 * 2.4 / 4


def fit_section(dev, num_jobs=4096, plain_per_class=2, rates=(200.0, 1000.0, 5000.0), tune=True,
                batched_only=False, cpu=True):
    """Card fitting: batched persistent launch, plain per-job loop, served
    continuous batching under Poisson arrivals (all 100 iterations per
    job), and the CPU loop of the oracle.  With tune, the launch shape is
    autotuned on the first 1024 jobs first; otherwise the library defaults
    are used.  batched_only stops after the persistent launch (profiler
    runs)."""
    import numpy as np
    import torch

    from strns2cards_amd import fitting

    t0 = time.perf_counter()
    jobs = fitting.make_jobs(num_jobs, seed=0)
    print(f"[bench] {num_jobs} fitting jobs generated in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    hyper = fitting.FitHyper()
    out = {}
    from strns2cards_amd import autotune

    if tune:
        tun = autotune.tune_fit(jobs[:1024], hyper, dev, iters=10)
        out["autotune"] = {"chunk": tun.chunk, "order": tun.order,
                           "ms_per_10_iters_1024_jobs": {k: round(v, 3) for k, v in tun.ms.items()}}
        print(f"[bench] fit autotune: {out['autotune']}", file=sys.stderr, flush=True)
    else:
        tun = autotune.FitTuning(0, "admission")

    # one persistent launch: every job runs all its iterations
    srv = fitting.CardFitServer(num_jobs, hyper, dev, chunk=tun.chunk, order=tun.order)
    slots = srv.admit(jobs)
    srv.run(0)
    torch.cuda.synchronize()
    init_loss = srv.loss.cpu().numpy()[slots]
    stream = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    srv.run(hyper.iters)
    b.record(stream)
    torch.cuda.synchronize()
    batched_s = a.elapsed_time(b) / 1e3
    final_loss = srv.loss.cpu().numpy()[slots]
    out["batched_jobs"] = num_jobs
    out["batched_launch_ms"] = round(batched_s * 1e3, 3)
    out["batched_iters_per_s"] = round(num_jobs * hyper.iters / batched_s, 1)
    print(f"[bench] batched: {num_jobs} jobs x {hyper.iters} its in {batched_s * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    prof = fit_roofline_profile()
    if prof and prof.get("jobs") == num_jobs and prof.get("iters") == hyper.iters:
        achieved = prof["valu_insts_per_launch"] / batched_s / 1e9
        out["roofline"] = {
            "bound": "valu", "achieved": round(achieved, 2), "peak": VALU_PEAK_GWIS,
            "unit": "G wave64 VALU instructions/s", "frac": round(achieved / VALU_PEAK_GWIS, 4),
            "traffic": prof.get("hbm_bytes_per_launch"), "kernel": "fit_cards_kernel",
            "avg_launch_ms": round(batched_s * 1e3, 3),
            "valu_insts_per_launch": prof["valu_insts_per_launch"],
            "source": "VALU count: profiles/fit_cards_sq_counters.md; time: this launch"}
    if batched_only:
        return out

    # plain per-job loop through the drop-in API, on every (resolution,
    # quads) class of the first 200 jobs, which every served rate completes
    sample = stratified(jobs[:200], plain_per_class)
    plain = {}
    fitting.plain_fit(jobs[0], hyper, dev, iters=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_tp = 0
    for j in sample:
        _, xor, tp = fitting.plain_fit(j, hyper, dev)
        plain[j.jid] = xor
        plain_tp += tp
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_ips = len(sample) * hyper.iters / plain_s
    plain_xor = sum(plain.values())
    batched_xor = float(sum(final_loss[j.jid] for j in sample))
    out["plain_jobs"] = len(sample)
    out["plain_iters_per_s"] = round(plain_ips, 2)
    out["plain_final_coverage_loss"] = round(plain_xor / max(plain_tp, 1), 5)
    out["batched_final_coverage_loss"] = round(batched_xor / max(plain_tp, 1), 5)
    out["batched_final_loss_rel_diff"] = round(abs(batched_xor - plain_xor) / max(plain_xor, 1), 6)
    out["initial_coverage_loss_all_jobs"] = float(init_loss.sum())
    out["final_coverage_loss_all_jobs"] = float(final_loss.sum())
    print(f"[bench] plain loop {plain_ips:.1f} it/s over {len(sample)} jobs, loss plain "
          f"{plain_xor} batched {batched_xor}", file=sys.stderr, flush=True)

    # served: Poisson arrivals, continuous batching, graph replays
    served = {}
    for lam in rates:
        n = min(num_jobs, max(200, int(lam * 1.0)))
        srv = fitting.CardFitServer(num_jobs, hyper, dev, chunk=tun.chunk, order=tun.order)
        # untimed warm-up on the same server (its graph, the allocators'
        # caches), as the DDA step gets warm-up steps
        fitting.serve(fitting.poisson_arrivals(jobs[-128:], lam, seed=1), srv, chunk=10)
        js = fitting.poisson_arrivals(jobs[:n], lam, seed=int(lam))
        rep = fitting.serve(js, srv, chunk=10)
        sxor = sum(rep.losses[j.jid] for j in sample)
        served[str(int(lam))] = {
            "jobs": n, "completed": rep.completed, "iters_per_s": round(rep.iters_per_s, 1),
            "p50_ms": round(rep.percentile(50), 3), "p99_ms": round(rep.percentile(99), 3),
            "makespan_s": round(rep.seconds, 3),
            "vs_plain": round(rep.iters_per_s / plain_ips, 1),
            "final_coverage_loss_plain_sample": round(sxor / max(plain_tp, 1), 5),
            "final_loss_rel_diff_vs_plain": round(abs(sxor - plain_xor) / max(plain_xor, 1), 6)}
        print(f"[bench] serve lambda={lam:.0f}: {served[str(int(lam))]}", file=sys.stderr, flush=True)
    out["served"] = served
    out["batched_vs_plain"] = round(out["batched_iters_per_s"] / plain_ips, 1)
    if cpu:
        c = fit_cpu_baseline(jobs, hyper)
        c["gpu_batched_vs_cpu"] = round(out["batched_iters_per_s"] / c["value"], 1)
        out["cpu_baseline"] = c
        print(f"[bench] card-fit cpu baseline {c}", file=sys.stderr, flush=True)
    return out


def adjoint_section(dev, steps=200):
    """Card-density fitting through the hand-written coverage adjoint: 1,024
    cards x 20 quads (centre-lines and widths as parameters) over 8 views at
    512 x 512, L2 against the 3x3 box-filtered DDA strand density, Adam at
    lr 1e-3."""
    import torch

    from strns2cards_amd import _core, adjoint, synth

    t0 = time.perf_counter()
    wl = synth.adjoint_workload()
    print(f"[bench] adjoint workload ready in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    pos = torch.from_numpy(wl.pos).to(dev)
    target = adjoint.density_target(pos, torch.from_numpy(wl.edges).to(dev),
                                    torch.from_numpy(wl.edge2cluster).to(dev), wl.num_clusters,
                                    (wl.H, wl.W))
    fit = adjoint.CardDensityFit.from_cards(wl.cards, wl.mvp, target, lr=1e-3)
    clip0 = fit.clip(fit.verts()).detach()
    fit.run(3)  # warm-up (allocator, Adam state); the timed run restarts from here
    rep = fit.run(steps)
    # forward / backward kernels alone at the fitted state, timed on their stream
    clip = fit.clip(fit.verts()).detach()
    stream = torch.cuda.current_stream()
    cov = torch.zeros((clip.shape[0], wl.H, wl.W), dtype=torch.float32, device=dev)
    d_pos = torch.zeros_like(clip)
    d_cov = torch.ones_like(cov)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = 20
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(reps):
        _core.card_coverage_forward(clip, fit.tri, cov)
    ev[1].record(stream)
    for _ in range(reps):
        _core.card_coverage_backward(clip, fit.tri, d_cov, d_pos)
    ev[2].record(stream)
    torch.cuda.synchronize()
    out = {"cards": int(wl.cards.tri.shape[0] // 40), "quads_per_card": 20, "views": int(clip.shape[0]),
           "resolution": [wl.H, wl.W], "steps": steps, "lr": 1e-3,
           "ms_per_step": round(rep.ms_per_step, 4),
           "steps_per_s": round(1e3 / rep.ms_per_step, 2),
           "loss_first": round(rep.loss_first, 6), "loss_last": round(rep.loss_last, 6),
           "forward_ms": round(ev[0].elapsed_time(ev[1]) / reps, 4),
           "backward_ms": round(ev[1].elapsed_time(ev[2]) / reps, 4),
           "initial_clip_vertices": int(clip0.shape[1])}
    print(f"[bench] adjoint fit: {out}", file=sys.stderr, flush=True)
    return out


def strip_layout_default(seed=0, guides=300, num_quads=15000):
    """Strip grids of 2_texture.py's defaults: `guides` guide strands with
    lognormal lengths share num_quads quads in proportion to length
    (allocate_vertex_counts), each swept into two crossed 2-row strips
    (res_w 2, tangent rotations -45 and 45 degrees)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    length = rng.lognormal(0.0, 0.5, size=guides)
    quads = np.maximum(1, np.rint(num_quads / 2 * length / length.sum())).astype(np.int64)
    cols = np.repeat(quads + 1, 2)
    return np.full_like(cols, 2), cols


def largesteps_section(dev, reps=50):
    """The largesteps system of 3_fitting.py:347-384 on grid meshes: factor
    once, then apply (to_differential) and solve (from_differential) timed
    with HIP events on the current stream, for the strips of 2_texture.py's
    defaults and for a batch of 16,384 strips of 2 x 128.  Algorithmic
    bytes: 24 per vertex (x or u read, the result written).  CPU
    counterpart: scipy's sparse LU of the same M (factored once), one solve
    of the default strips on one core."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from strns2cards_amd.optimizer import GridSystem, grid_edges

    stream = torch.cuda.current_stream()
    out = {}
    for name, (rows, cols) in [("reference_default_strips", strip_layout_default()),
                               ("batch_16384x2x128", (np.full(16384, 2), np.full(16384, 128)))]:
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record(stream)
        M = GridSystem(rows, cols, 100.0, 0.01, device=dev)
        ev[1].record(stream)
        g = torch.Generator(device=dev).manual_seed(1)
        x = torch.randn(M.num_verts, 3, device=dev, generator=g)
        u = M.apply(x)
        M.solve(u)
        torch.cuda.synchronize()
        ev[2].record(stream)
        for _ in range(reps):
            M.apply(x)
        ev[3].record(stream)
        for _ in range(reps):
            M.solve(u)
        ev[4].record(stream)
        torch.cuda.synchronize()
        apply_ms = ev[2].elapsed_time(ev[3]) / reps
        solve_ms = ev[3].elapsed_time(ev[4]) / reps
        nbytes = 24 * M.num_verts
        err = ((M.solve(u) - x).norm() / x.norm()).item()
        out[name] = {"grids": int(M.num_grids), "vertices": int(M.num_verts),
                     "max_columns": int(np.max(cols)),
                     "factor_ms_incl_setup": round(ev[0].elapsed_time(ev[1]), 4),
                     "apply_us": round(apply_ms * 1e3, 2), "solve_us": round(solve_ms * 1e3, 2),
                     "solve_GBps": round(nbytes / solve_ms / 1e6, 1),
                     "solve_frac_of_hbm": round(nbytes / solve_ms / 1e6 / HBM_PEAK_GBS, 4),
                     "round_trip_rel_err": float(f"{err:.3e}")}
        if name == "reference_default_strips":
            V = M.num_verts
            eh, ew = grid_edges(M.rows_host, M.cols_host)

            def lap(e):
                e = np.concatenate([e, e[:, ::-1]])
                A = sp.coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(V, V)).tocsr()
                return sp.diags(np.asarray(A.sum(1)).ravel()) - A

            Ms = (sp.eye(V) + 100.0 * (lap(eh) + 0.01 * lap(ew))).tocsc()
            t0 = time.perf_counter()
            lu = spla.splu(Ms)
            t1 = time.perf_counter()
            ub = u.cpu().numpy().astype(np.float64)
            n = 0
            while time.perf_counter() - t1 < 1.0 or n < 3:
                lu.solve(ub)
                n += 1
            t2 = time.perf_counter()
            out[name]["cpu_scipy_splu"] = {"factor_ms": round((t1 - t0) * 1e3, 3),
                                           "solve_us": round((t2 - t1) / n * 1e6, 1),
                                           "cores": 1, "dtype": "f64"}
    print(f"[bench] largesteps: {out}", file=sys.stderr, flush=True)
    return out


def card_mesh_section(dev, reps=20):
    """Card mesh inputs on the GPU for 2_texture.py's defaults (300 guide
    strands of lognormal lengths sharing 7,500 quads, two crossed 2-row
    strips each): strips (strand_to_strip_mesh), then the grid meshes'
    faces / quads / uvs / h_edges / w_edges / face clusters
    (combine_wisps_with_cluster_labels), timed with HIP events.  CPU
    baseline: the oracle's numpy restatement of the same two functions
    (the reference computes them with numpy on the host), one core."""
    import numpy as np
    import torch

    import oracle
    from strns2cards_amd import cardmesh

    rng = np.random.default_rng(3)
    rows, cols = strip_layout_default()
    counts = cols[::2]
    strands = [np.cumsum(rng.normal(size=(n, 3)) * 0.01, 0).astype(np.float32) for n in counts]
    rects = rng.uniform(0, 1, (len(cols), 4))
    g2c = rng.integers(0, 300, len(cols))
    pts, cnt, offs, _ = cardmesh._ragged(strands, dev)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    cardmesh.strands_to_strips_device(pts, cnt, offs, 2, 0.1, [-45.0, 45.0])
    cardmesh.grid_mesh_topology_device(rows, cols, rects, g2c, device=dev)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(reps):
        cardmesh.strands_to_strips_device(pts, cnt, offs, 2, 0.1, [-45.0, 45.0])
    ev[1].record(stream)
    for _ in range(reps):
        cardmesh.grid_mesh_topology_device(rows, cols, rects, g2c, device=dev)
    ev[2].record(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in strands:
        oracle.strand_to_strip_mesh(s, 2, 0.1, [-45.0, 45.0])
    t1 = time.perf_counter()
    oracle.combine_grid_meshes(list(zip(rows.tolist(), cols.tolist())), rects.tolist(),
                               g2c.tolist())
    t2 = time.perf_counter()
    out = {"guides": len(strands), "strips": int(len(cols)), "vertices": int(2 * cols.sum()),
           "quads": int(((rows - 1) * (cols - 1)).sum()),
           "strips_ms": round(ev[0].elapsed_time(ev[1]) / reps, 4),
           "topology_ms_incl_host_layout": round(ev[1].elapsed_time(ev[2]) / reps, 4),
           "cpu_baseline": {"strips_ms": round((t1 - t0) * 1e3, 2),
                            "topology_ms": round((t2 - t1) * 1e3, 2), "cores": 1,
                            "kind": "port (numpy, the reference's host algorithm)"}}
    print(f"[bench] card mesh: {out}", file=sys.stderr, flush=True)
    return out


def texture_section(dev, reps=5):
    """One wisp's software-MSAA texture patch (render_texture,
    texture_utils.py:203-254): a 200-strand x 64-point wisp as 8-gon tubes
    (102,400 vertices, 201,600 triangles) into a 256 x 128 patch, 64
    jitters.  Fused: one 64-view depth-key pass + resolve + blur
    (texture.render_patch); batched: the same through one 64-view drop-in
    rasterize (fragment lists).  Reference-shaped: the 64 single-view
    rasterize + interpolate calls and torch accumulation of the reference's
    loop, on the same drop-in kernels.  CPU baseline: the oracle's C
    rasterizer and numpy accumulation on 4 of the 64 jitters, scaled."""
    import numpy as np
    import torch

    import oracle
    from strns2cards_amd import texture
    from strns2cards_amd.rasterize import interpolate_triangle_attributes, rasterize

    rng = np.random.default_rng(5)
    base = np.cumsum(rng.normal(size=(64, 3)) * [0.01, 0.05, 0.01] + [0, 0.05, 0], 0)
    wisp = np.stack([base + rng.normal(size=3) * [0.05, 0.01, 0.05] for _ in range(200)])
    wisp = wisp.astype(np.float32)
    p = wisp.reshape(-1, 3)
    bbox = p[:, :2].max(0) - p[:, :2].min(0)
    scale = 2.0 * np.ones(3) / (p.max(0) - p.min(0))
    scale[2] *= 0.2
    verts, faces, attr = texture.wisp_ndc_mesh(wisp, (0.0, 0.0, 0.5, 1.0), bbox, scale, 0.002, 8)
    verts[:, 0] *= 0.9  # keep the strands inside the patch
    H, W = 256, 128
    pos = torch.from_numpy(verts.astype(np.float32)).to(dev)
    tri = torch.from_numpy(faces).to(dev)
    at = torch.from_numpy(np.ascontiguousarray(attr, np.float32)).to(dev)
    offs, wts = texture.msaa_samples()

    def batched():
        return texture.render_patch(pos, tri, at, (H, W))

    def batched_raster():
        return texture.render_patch(pos, tri, at, (H, W), fused=False)

    def looped():
        acc = torch.zeros((H, W, 3), device=dev)
        cnt = torch.zeros((H, W), device=dev)
        acc_alpha = 0
        for i, (ox, oy) in enumerate(offs):
            pj = pos.clone()
            pj[:, 0] += ox * 2.0 / W
            pj[:, 1] += oy * 2.0 / H
            ph = torch.cat([pj, torch.ones_like(pos[:, :1])], -1)
            rast = rasterize((H, W), ph[None], tri)
            rgb = interpolate_triangle_attributes(rast, at, tri)[0]
            a = (rast[0, ..., -1] > 0).float()
            acc += rgb * a[..., None]
            cnt += a
            acc_alpha = acc_alpha + wts[i] * a
        return acc, cnt, acc_alpha

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    ms_b, ms_r, ms_l = timed(batched), timed(batched_raster), timed(looped)
    dx = np.array([o[0] * 2.0 / W for o in offs], np.float32)
    dy = np.array([o[1] * 2.0 / H for o in offs], np.float32)
    t0 = time.perf_counter()
    oracle.msaa_patch(verts, faces, attr, H, W, dx[:4], dy[:4], np.array(wts[:4], np.float32),
                      texture.gaussian_kernel3(1.5))
    cpu_ms = (time.perf_counter() - t0) * 1e3 * 16
    out = {"strands": 200, "points": 64, "tube_sides": 8, "vertices": int(verts.shape[0]),
           "triangles": int(faces.shape[0]), "patch": [H, W], "samples": 64,
           "fused_ms": round(ms_b, 3), "batched_rasterize_ms": round(ms_r, 3),
           "per_jitter_loop_ms": round(ms_l, 3), "fused_vs_loop": round(ms_l / ms_b, 2),
           "cpu_baseline": {"ms": round(cpu_ms, 1), "cores": 1,
                            "kind": "port (C raster oracle + numpy)",
                            "sample": "4 of 64 jitters, x16"}}
    # wisp masks of render_patches_reduced: 300 wisps of 50 strands x 64
    # points into 256 x 128 patches, one launch against the reference's
    # per-wisp dda_compute_fragments + depth_test + dda_interpolate loop
    from strns2cards_amd.dda import dda_compute_fragments, dda_interpolate_attributes
    from strns2cards_amd.rasterize import depth_test

    wisps = []
    for _ in range(300):
        b = np.cumsum(rng.normal(size=(64, 3)) * [0.01, 0.05, 0.01] + [0, 0.05, 0], 0)
        wisps.append(np.stack([b + rng.normal(size=3) * [0.05, 0.01, 0.05]
                               for _ in range(50)]).astype(np.float32))
    uv = [(i, 0.0, 0.0, 0.5, 1.0) for i in range(len(wisps))]
    Rm = 256
    t0 = time.perf_counter()
    masks = texture.patch_masks(wisps, uv, Rm, device=dev)
    torch.cuda.synchronize()
    ms_masks = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    masks = texture.patch_masks(wisps, uv, Rm, device=dev)
    torch.cuda.synchronize()
    ms_masks = min(ms_masks, (time.perf_counter() - t0) * 1e3)
    Hm, Wm = masks.shape[1:]
    # the one launch alone, on device-resident points
    _, _, pbm, scm = texture.packing_frame(wisps)
    posm = torch.ones((len(wisps) * 50 * 64, 4), device=dev)
    posm[:, :3] = torch.from_numpy(np.concatenate(
        [texture.wisp_ndc_points(w_, r[1:5], pbm, scm) for w_, r in zip(wisps, uv)])).to(dev)
    evm = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    texture.patch_masks_device(posm, [50] * len(wisps), [64] * len(wisps), (Hm, Wm))
    torch.cuda.synchronize()
    evm[0].record(torch.cuda.current_stream())
    for _ in range(reps):
        texture.patch_masks_device(posm, [50] * len(wisps), [64] * len(wisps), (Hm, Wm))
    evm[1].record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    ms_dev = evm[0].elapsed_time(evm[1]) / reps
    allp = np.concatenate([w.reshape(-1, 3) for w in wisps])
    bb = allp[:, :2].max(0) - allp[:, :2].min(0)
    sz = np.zeros(3)
    for w_ in wisps:
        q = w_.reshape(-1, 3)
        sz = np.maximum(sz, q.max(0) - q.min(0))
    sc = 2.0 * np.ones(3) / sz
    sc[2] *= 0.2
    idx = np.arange(50 * 64).reshape(50, 64)
    edges_np = np.stack([idx[:, :-1].reshape(-1), idx[:, 1:].reshape(-1)], 1).astype(np.int32)
    edges_t = torch.from_numpy(edges_np).to(dev)
    loop_pos = []
    for w_, r in zip(wisps, uv):
        ph = torch.ones((50 * 64, 4), device=dev)
        ph[:, :3] = torch.from_numpy(texture.wisp_ndc_points(w_, r[1:5], bb, sc)).to(dev)
        loop_pos.append(ph[None])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ph in loop_pos:
        fp, fa = dda_compute_fragments((Hm, Wm), ph, edges_t)
        rast = depth_test(1, (Hm, Wm), fp, fa)
        mk = dda_interpolate_attributes(rast, torch.ones_like(ph[0, :, :1]), edges_t)[0]
        mk = (torch.flip(mk[..., 0], dims=(0,)).clamp(0, 1) * 255).to(torch.uint8)
    torch.cuda.synchronize()
    ms_loop = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for w_, r in zip(wisps[:5], uv[:5]):
        oracle.patch_mask(texture.wisp_ndc_points(w_, r[1:5], bb, sc), 50, 64, Hm, Wm)
    cpu_masks = (time.perf_counter() - t0) * 1e3 * len(wisps) / 5
    out["patch_masks"] = {"wisps": len(wisps), "strands_per_wisp": 50, "points": 64,
                          "patch": [int(Hm), int(Wm)],
                          "one_launch_ms": round(ms_dev, 3),
                          "with_host_prep_and_copy_ms": round(ms_masks, 2),
                          "per_wisp_loop_ms": round(ms_loop, 2),
                          "cpu_baseline": {"ms": round(cpu_masks, 1), "cores": 1,
                                           "kind": "port (C DDA oracle + numpy)",
                                           "sample": "5 of 300 wisps, scaled"}}
    print(f"[bench] texture patch: {out}", file=sys.stderr, flush=True)
    return out


def joint_fit_section(dev, iters=50, fused_only=False):
    """3_fitting.py's loop at its defaults (13 views at 200 x 200, 300
    clusters, about 15,000 card quads in 600 crossed 2-row strips, large
    steps lam 100 / w 1/100, lr 0.1) on synthetic strands (10,000 x 64
    points): ms per iteration through the drop-in API as the script calls
    it, and with `fused=True` (triangle_cluster_mask, the grid kernels, the
    antialias gradient from the marked edges and the two-launch UniformAdam
    step, no host sync), and the XOR
    count over all views before and after `iters` iterations.  `fused_only`
    skips the script-shaped loop (for kernel traces of the fused one)."""
    from strns2cards_amd import fitting

    t0 = time.perf_counter()
    inp = fitting.joint_fit_inputs(device=dev)
    prep = time.perf_counter() - t0
    if fused_only:
        fitting.joint_fit(**inp, num_iters=3, fused=True)
        rep_f = fitting.joint_fit(**inp, num_iters=iters, fused=True)
        out = {"iters": iters, "fused_ms_per_iter": round(rep_f.ms_per_iter, 3),
               "xor_first": rep_f.xor_first, "fused_xor_last": rep_f.xor_last}
        print(f"[bench] joint fit (fused only): {out}", file=sys.stderr, flush=True)
        return out
    fitting.joint_fit(**inp, num_iters=3)  # warm-up
    rep = fitting.joint_fit(**inp, num_iters=iters)
    fitting.joint_fit(**inp, num_iters=3, fused=True)
    rep_f = fitting.joint_fit(**inp, num_iters=iters, fused=True)
    out = {"strands": 10000, "points": 64, "clusters": 300, "views": 13, "resolution": 200,
           "card_vertices": int(inp["card_pos"].shape[0]), "quads": int(inp["quads"].shape[0]),
           "iters": iters, "ms_per_iter": round(rep.ms_per_iter, 3),
           "iters_per_s": round(1e3 / rep.ms_per_iter, 1),
           "xor_first": rep.xor_first, "xor_last": rep.xor_last,
           "fused_ms_per_iter": round(rep_f.ms_per_iter, 3),
           "fused_iters_per_s": round(1e3 / rep_f.ms_per_iter, 1),
           "fused_xor_last": rep_f.xor_last,
           "input_prep_s": round(prep, 2)}
    print(f"[bench] joint fit: {out}", file=sys.stderr, flush=True)
    return out


def copy_calibration(dev, nbytes=1 << 29, reps=3):
    """A plain device-to-device copy of known size, so PMC byte counters can
    be scaled against it."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: {reps} x {nbytes} bytes", file=sys.stderr, flush=True)


def api_form_section(s2c, w, pos, edges, e2c, reps=5):
    """The drop-in API form at the full size: fragment lists (dda.py:5-76),
    then the mask from them (rasterize.py:170-214), as 3_fitting.py:254-261
    calls them.  Wall-clock per call (each holds host synchronizations)."""
    import torch

    res = (w.H, w.W)
    B, V, _ = w.pos.shape
    E = w.edges.shape[0]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    fp, fa = s2c.dda_compute_fragments(res, pos, edges)
    nf = int(fp.shape[0])
    del fp, fa
    ms_api = timed(lambda: s2c.dda_compute_fragments(res, pos, edges))
    ms_api_mask = timed(lambda: s2c.cluster_mask_from_fragments(
        res, *s2c.dda_compute_fragments(res, pos, edges), e2c, w.num_clusters))
    api_bytes = B * V * 16 + E * 8 + nf * 28
    print(f"[bench] API form: dda_compute_fragments {ms_api:.3f} ms ({nf} fragments, "
          f"{api_bytes / ms_api / 1e6:.0f} GB/s), + cluster_mask_from_fragments {ms_api_mask:.3f} ms",
          file=sys.stderr, flush=True)
    return {"dda_api_fragments_ms": round(ms_api, 4), "dda_api_fragments": nf,
            "dda_api_algorithmic_bytes": api_bytes,
            "dda_api_GBps": round(api_bytes / ms_api / 1e6, 1),
            "dda_api_fragments_plus_mask_ms": round(ms_api_mask, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--fit-jobs", type=int, default=4096)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-largesteps", action="store_true",
                    help="skip the large-steps, card-mesh and texture sections")
    ap.add_argument("--largesteps-only", action="store_true",
                    help="only the large-steps, card-mesh, texture and joint-fit sections")
    ap.add_argument("--joint-fit-only", action="store_true",
                    help="only the fused joint-fit loop (for kernel traces)")
    ap.add_argument("--fit-only", action="store_true",
                    help="only the persistent card-fitting launch (for PMC runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the coverage-adjoint fit section (for PMC runs)")
    ap.add_argument("--cards-only", action="store_true",
                    help="stop after the card raster and antialias timings (for PMC runs)")
    ap.add_argument("--card-reps", type=int, default=10,
                    help="timed repetitions of each card stage")
    ap.add_argument("--api-only", action="store_true",
                    help="only the drop-in API form (dda_compute_fragments, "
                         "cluster_mask_from_fragments) at the full size")
    ap.add_argument("--dda-only", action="store_true",
                    help="time the DDA->mask step only (for PMC runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuners (library defaults; for profiler runs)")
    ap.add_argument("--dda-tile", type=int, default=0,
                    help="DDA->mask tile side log2 (implies no DDA autotune; for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 512 MiB device copy (PMC byte-counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the path's multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    import torch

    import strns2cards_amd as s2c
    from strns2cards_amd import _core, synth

    dev = torch.device("cuda", 0)
    if args.fit_only or args.adjoint_only:
        if args.fit_only:
            res = {"card_fit": fit_section(dev, num_jobs=args.fit_jobs, tune=not args.no_autotune,
                                           batched_only=True)}
        else:
            res = {"adjoint_fit": adjoint_section(dev)}
        if args.copy_calibration:
            copy_calibration(dev)
        print(json.dumps(res), flush=True)
        return
    if args.joint_fit_only:
        print(json.dumps({"joint_fit": joint_fit_section(dev, fused_only=True)}), flush=True)
        return
    if args.largesteps_only:
        print(json.dumps({"largesteps": largesteps_section(dev),
                          "card_mesh": card_mesh_section(dev),
                          "texture_patch": texture_section(dev),
                          "joint_fit": joint_fit_section(dev)}), flush=True)
        return
    t_gen = time.perf_counter()
    w = synth.full_workload()
    gen_s = time.perf_counter() - t_gen
    print(f"[bench] workload ready in {gen_s:.1f}s: {w.name}", file=sys.stderr, flush=True)

    B, V, _ = w.pos.shape
    E = w.edges.shape[0]
    S = (w.num_clusters + 31) // 32
    pos = torch.from_numpy(w.pos).to(dev)
    edges = torch.from_numpy(w.edges).to(dev)
    e2c = torch.from_numpy(w.edge2cluster).to(dev)
    bitset = torch.empty((B, w.H, w.W, S), dtype=torch.int32, device=dev)

    if args.api_only:
        print(json.dumps(api_form_section(s2c, w, pos, edges, e2c)), flush=True)
        return

    from strns2cards_amd import autotune

    if args.no_autotune or args.dda_tile:
        dda_tune = autotune.DdaTuning(args.dda_tile)
    else:
        dda_tune = autotune.tune_dda(pos, edges, e2c, bitset)
        print(f"[bench] dda tile autotune: {dda_tune.ms} -> side {1 << dda_tune.tile_log2}",
              file=sys.stderr, flush=True)

    def step():
        _core.dda_cluster_mask(pos, edges, e2c, bitset, dda_tune.tile_log2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    # whole-job rate: K steps back to back between two events
    torch.cuda.synchronize()
    t_all0.record(stream)
    for _ in range(args.steps):
        step()
    t_all1.record(stream)
    torch.cuda.synchronize()
    ms_total = t_all0.elapsed_time(t_all1)
    ms_step = ms_total / args.steps
    # average launch duration: the same K steps, each between its own events
    # (the events' own packets stay out of the rate above)
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    per_launch = sorted(a.elapsed_time(b) for a, b in evs)
    ms_launch = sum(per_launch) / len(per_launch)
    nbytes = dda_mask_bytes(B, V, E, w.H, w.W, S)
    value = nbytes / (ms_step / 1e3) / 1e9
    achieved = nbytes / (ms_launch / 1e3) / 1e9
    print(f"[bench] dda->mask {ms_step:.3f} ms/step, {value:.1f} GB/s", file=sys.stderr, flush=True)

    if args.copy_calibration and not args.cards_only:
        copy_calibration(dev)
    if args.dda_only:
        print(json.dumps({"metric": METRIC, "value": round(value, 3),
                          "unit": "GB/s (DDA->mask algorithmic bytes)", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(ms_step, 4), "dda_only": True}), flush=True)
        return

    # Card stages at the same size (reported beside the headline).
    c = w.cards
    cpos = torch.from_numpy(w.card_pos).to(dev)
    tri = torch.from_numpy(c.tri).to(dev)
    t2c = torch.from_numpy(c.tri2cluster).to(dev)
    cedges = torch.from_numpy(c.edges).to(dev)
    e2t = torch.from_numpy(c.edge2tri).to(dev)
    ce2c = torch.from_numpy(c.edge2cluster).to(dev)
    cbits = torch.empty_like(bitset)

    def timed(fn, reps=args.card_reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    ms_tri = timed(lambda: _core.triangle_cluster_mask(cpos, tri, t2c, cbits))
    ms_aa = timed(lambda: s2c.antialiased_cluster_bitset_loss(
        cbits, bitset, cpos, tri, cedges, e2t, ce2c, 7.0, 0.5, False))
    ms_aa_f = timed(lambda: s2c.antialiased_cluster_bitset_loss(
        cbits, bitset, cpos, tri, cedges, e2t, ce2c, 7.0, 0.5, False, fused=True))
    print(f"[bench] card raster->mask {ms_tri:.3f} ms, antialias grad {ms_aa:.3f} ms "
          f"(from the marked edges, no host sync: {ms_aa_f:.3f} ms)", file=sys.stderr, flush=True)
    if args.cards_only:
        if args.copy_calibration:
            copy_calibration(dev)
        print(json.dumps({"card_raster_mask_ms": round(ms_tri, 4),
                          "antialias_grad_ms": round(ms_aa, 4),
                          "antialias_grad_fused_ms": round(ms_aa_f, 4)}), flush=True)
        return
    api = api_form_section(s2c, w, pos, edges, e2c)

    fit = None if args.no_fit else fit_section(dev, num_jobs=args.fit_jobs,
                                                   tune=not args.no_autotune, cpu=not args.no_cpu)
    adj = None if args.no_adjoint else adjoint_section(dev)
    lsteps = None if args.no_largesteps else largesteps_section(dev)
    cmesh = None if args.no_largesteps else card_mesh_section(dev)
    tex = None if args.no_largesteps else texture_section(dev)
    jfit = None if args.no_fit else joint_fit_section(dev)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_seconds)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s (DDA->mask algorithmic bytes)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded strands/cards; no dataset)",
        "config": {"workload": w.name, "views": B, "resolution": [w.H, w.W],
                   "strands": int(w.strands.shape[0]), "vertices_per_strand": int(w.strands.shape[1]),
                   "clusters": w.num_clusters, "bitset_words": S, "edges": E,
                   "algorithmic_bytes_per_step": nbytes,
                   "autotune": None if args.no_autotune or args.dda_tile else
                   {"tile_side": 1 << dda_tune.tile_log2,
                    "ms_per_tile_side": {str(k): round(v, 4) for k, v in dda_tune.ms.items()}}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic_from_profiles(),
                     "kernel": "s2c_dda_cluster_mask entry point (dda_cell_stage_kernel, "
                               "bin_count/colsum/colprefix/scan/order/place kernels, "
                               "cell_resolve_kernel, dda_mask_fallback_kernel)",
                     "avg_launch_ms": round(ms_launch, 4)},
        "cpu_baseline": cpu,
        "extra": {"card_raster_mask_ms": round(ms_tri, 4), "antialias_grad_ms": round(ms_aa, 4),
                  "antialias_grad_fused_ms": round(ms_aa_f, 4),
                  **api,
                  "cards": int(c.tri.shape[0] // 40), "card_triangles": int(c.tri.shape[0])},
        "card_fit": fit,
        "adjoint_fit": adj,
        "largesteps": lsteps,
        "card_mesh": cmesh,
        "texture_patch": tex,
        "joint_fit": jfit,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
