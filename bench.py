#!/usr/bin/env python3
"""Benchmark: B12's set-oriented hot path on one MI355X.

Workload (one "step" = one whole job of the Lorenz configuration): Lorenz flow map
(sigma=10, rho=28, beta=8/3) as 20 RK4 steps of h=0.01, FP64; domain centre (0,0,27), radius
(30,30,30); 24 subdivision steps {subdivide(SD); setFlagsByIterators(OverLeaves(InnerGrid(3)),
HIT); remove(HIT); unsetFlags(HIT)}; then the transition matrix on the final cover and the
invariant measure by power iteration until ||mu_{k+1} - mu_k||_1 <= 1e-12.

Metric: test-point images per second over the whole job (images of the 24 subdivision levels
plus the transition-matrix pass). Inputs are generated on the GPU; nothing crosses PCIe in the
timed region except the per-level counts the API returns.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6          # MI355X FP64 vector, FMA counted as 2 flops
PEAK_FP64_ISSUE_TFLOPS = 39.3    # the same peak when every FP64 op is a separate instruction
CUS = 256                        # MI355X compute units (4 SIMDs of 16 FP64 lanes each)
LORENZ_FLOPS_PER_IMAGE = 1590    # 20 RK4 steps x 78 + grid/affine/key 30 (each + - * / once)

PEAK_HBM_BYTES = 8.0e12          # MI355X HBM3E
# algorithmic bytes per leaf and subdivision step: subdivide (10 B read, 20 B written per parent,
# i.e. 15 B per child), image + flag (10 B of the box, 1 B hit store per test point), remove +
# unset (11 B read, 10 B written)
BYTES_PER_BOX_STEP = 15 + 10 + 27 + 21
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "lorenz_d24_job_traffic.md")

CENTER, RADIUS, DEPTH, K = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], 24, 27
TOL = 1e-12
MAX_ITER = 1_000_000


def images_of(counts, n0=1):
    """Test points imaged by the subdivision steps: k x (leaves after each subdivide)."""
    imgs, n = 0, n0
    for c in counts:
        n_sub = n + int(c[0])   # every flagged leaf splits in two (depth < 64)
        imgs += K * n_sub
        n = int(c[3])           # leaves kept (all carried HIT)
    return imgs, n


def profiled_traffic(kernel="imageKernel<3,LORENZ,INNER>"):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE profile (scaled
    by the copy-kernel calibration), or None."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\(" + re.escape(kernel) + r"\):\s*(\d+)", text)
    return int(m.group(1)) if m else None


def gpu_job(gx, flavour, timing=False):
    t = gx.ImplicitBoxTree(CENTER, RADIUS, flavour=flavour)
    grid, mp = gx.InnerGrid(3), gx.Lorenz()
    if timing:
        t.image_timing(1)
    t.setFlags(gx.SD, 0)
    t0 = time.perf_counter()
    counts = t.subdivision_steps(grid, mp, DEPTH)
    t1 = time.perf_counter()
    m = t.transitionMatrixForLeaves(grid, mp)
    t2 = time.perf_counter()
    mu, iters, resid = m.power_iteration(MAX_ITER, TOL)
    t3 = time.perf_counter()
    pi_ms = gx.api.power_iteration_kernel_ms()
    imgs, n = images_of(counts)
    imgs += K * n
    res = {"images": imgs, "boxes": n, "nnz": m.nnz, "iters": iters, "resid": resid, "pi_kernel_ms": pi_ms,
           "boxes_stepped": (imgs - K * n) // K,
           "subdiv_ms": (t1 - t0) * 1e3, "tm_ms": (t2 - t1) * 1e3, "measure_ms": (t3 - t2) * 1e3,
           "level24_points": K * (int(counts[-2][3]) + int(counts[-1][0])) if DEPTH > 1 else 0}
    if timing:
        res["image_launches"] = t.image_launches()
        res["image_timing"] = t.image_timing(0)
    return res


def measure_roofline(job):
    """The job's dominant kernel, the persistent power iteration (one launch per solve): its
    algorithmic bytes are 12 per nonzero (row index, value) and 16 per column (column offset,
    z read, z written, the norm's share) per iteration; HIP events on the matrix's stream time
    the launch."""
    n, nnz, iters, ms = job["boxes"], job["nnz"], job["iters"], job["pi_kernel_ms"]
    alg = iters * (12 * nnz + 16 * n)
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": "piFlowKernel<false>", "achieved": achieved, "peak": PEAK_HBM_BYTES / 1e9,
            "unit": "GB/s", "frac": achieved * 1e9 / PEAK_HBM_BYTES, "traffic": profiled_traffic("piFlowKernel"),
            "traffic_source": "profiles/lorenz_d24_job_traffic.md (FETCH_SIZE x2.000 + WRITE_SIZE x1.000)",
            "algorithmic_bytes": alg, "launch_ms": ms, "iterations": iters,
            "us_per_iteration": ms * 1e3 / iters if iters else None,
            "note": "latency-bound: each iteration hands z_t from its producers to its consumers through "
                    "the memory-side cache (a store, then polled gathers), about 3.7 us per iteration"}


HENON_C, HENON_R = [0.0, 0.0], [1.5, 0.5]


def sweep_jobs(n_jobs, seed=0):
    """Config 4: Hénon a ~ U[1.0, 1.4], b = 0.3, depth ~ U{14..22} (numpy PCG64, seed 0, drawn
    in the order (a, depth) per job)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(1.0, 1.4)), int(rng.integers(14, 23))) for _ in range(n_jobs)]


def plain_henon_job(gx, grid, a, d, pi_iter):
    t = gx.ImplicitBoxTree(HENON_C, HENON_R)
    t.setFlags(gx.SD, 0)
    t.subdivision_steps(grid, gx.Henon(a, 0.3, 1), d)
    m = t.transitionMatrixForLeaves(grid, gx.Henon(a, 0.3, 1))
    mu, it, _ = m.power_iteration(pi_iter)
    return t, m, mu, it


def sweep_verify(gx, jobs, grid, pi_iter):
    """Untimed check before the timed runs: every job of the batched form (one forest) has the
    cover, matrix, measure and iteration count of its plain single-tree run, bit for bit."""
    import numpy as np
    F = gx.Forest(2, grid, 1, pi_max_iter=pi_iter)
    ids = [F.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d) for a, d in jobs]
    F.run()
    boxes = []
    for j, (a, d) in zip(ids, jobs):
        t, m, mu, it = plain_henon_job(gx, grid, a, d, pi_iter)
        same = (F.info(j)["iters"] == it and all(np.array_equal(x, y) for x, y in zip(F.cover(j), t.download()))
                and all(np.array_equal(x, y) for x, y in zip(F.matrix(j), m.download()))
                and np.array_equal(F.measure(j), mu))
        if not same:
            raise AssertionError(f"sweep job {j} (a={a}, depth {d}): batched result differs from the plain run")
        boxes.append(t.size())
    return min(boxes), max(boxes)


def sweep_bench(gx, n_jobs, pi_iter=200):
    jobs = sweep_jobs(n_jobs)
    grid = gx.InnerGrid(4)
    lo, hi = sweep_verify(gx, jobs, grid, pi_iter)
    # plain form: one job per launch sequence, sequentially
    t0 = time.perf_counter()
    for a, d in jobs:
        t = gx.ImplicitBoxTree(HENON_C, HENON_R)
        t.setFlags(gx.SD, 0)
        t.subdivision_steps(grid, gx.Henon(a, 0.3, 1), d)
        m = t.transitionMatrixForLeaves(grid, gx.Henon(a, 0.3, 1))
        m.power_iteration(pi_iter)
    plain_s = time.perf_counter() - t0
    # batched form: every job in one forest, shared launches
    t0 = time.perf_counter()
    F = gx.Forest(2, grid, 1, pi_max_iter=pi_iter)
    ids = [F.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d) for a, d in jobs]
    F.run()
    batched_s = time.perf_counter() - t0
    assert all(F.info(i)["state"] == 2 for i in ids)
    return {"config": f"{n_jobs} Hénon jobs, a~U[1.0,1.4], b=0.3, depth~U{{14..22}}, k=16, TM + {pi_iter} power "
                      f"iterations each", "plain_jobs_per_s": n_jobs / plain_s,
            "batched_jobs_per_s": n_jobs / batched_s, "ratio": plain_s / batched_s,
            "plain_s": plain_s, "batched_s": batched_s, "covers_boxes": [lo, hi],
            "verified": "every job's cover, matrix, measure and iteration count equal its plain run (bit for bit)"}


def serve_requests(n_req, rate=200.0, seed=0):
    """Config 5: Poisson arrivals at `rate`; 60% Lorenz (rho~U[24,32], depth U{18..26}, k=27,
    RK4 h=0.01, T=0.2), 40% forced Duffing (period map, 100 RK4 steps, domain radius (2,2),
    depth U{16..24}, k=16)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n_req))
    req = []
    for i in range(n_req):
        if rng.random() < 0.6:
            req.append((0, [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], [10.0, float(rng.uniform(24, 32)), 8.0 / 3.0, 0.01, 20.0],
                        int(rng.integers(18, 27))))
        else:
            req.append((1, [0.0, 0.0], [2.0, 2.0], [0.25, -1.0, 1.0, 0.3, 1.0, 0.0, 2 * 3.141592653589793,
                                                   2 * 3.141592653589793 / 100], int(rng.integers(16, 25))))
    return arrival, req


def serve_bench(gx, n_req, rate=200.0, tol=TOL, n_steady=0):
    import numpy as np
    arrival, req = serve_requests(n_req, rate)
    def forests():
        return [gx.Forest(3, gx.InnerGrid(3), 2, pi_max_iter=MAX_ITER, pi_tol=tol),
                gx.Forest(2, gx.InnerGrid(4), 4, pi_max_iter=MAX_ITER, pi_tol=tol)]
    def request_args(req):
        return ([r[0] for r in req], [r[1] for r in req], [r[2] for r in req], [r[3] for r in req], [r[4] for r in req])
    args = request_args(req)
    fs = forests()
    lat, comp, _, info = gx.serve_run(fs, arrival, *args)
    served = {"jobs_per_s": n_req / comp.max(), "p50_s": float(np.percentile(lat, 50)),
              "p99_s": float(np.percentile(lat, 99)), "makespan_s": float(comp.max()),
              "peak_footprint_bytes": info["peak_footprint"], "max_in_flight": info["max_in_flight"],
              "deferred_by_budget": info["deferred"]}
    fs = forests()
    lat0, comp0, _, _ = gx.serve_run(fs, np.zeros(n_req), *args)
    saturated = {"jobs_per_s": n_req / comp0.max(), "p50_s": float(np.percentile(lat0, 50)),
                 "p99_s": float(np.percentile(lat0, 99))}
    # plain: the same jobs one at a time through the single-tree path, in arrival order
    t0 = time.perf_counter()
    plain_lat = []
    clock = 0.0
    for (q, c, r, mp, d), a in zip(req, arrival):
        s0 = time.perf_counter()
        grid = gx.InnerGrid(3) if q == 0 else gx.InnerGrid(4)
        mapf = gx.Lorenz(*mp) if q == 0 else gx.Duffing(*mp)
        t = gx.ImplicitBoxTree(c, r)
        t.setFlags(gx.SD, 0)
        t.subdivision_steps(grid, mapf, d)
        m = t.transitionMatrixForLeaves(grid, mapf)
        m.power_iteration(MAX_ITER, tol)
        dur = time.perf_counter() - s0
        clock = max(clock, a) + dur      # a single server queue
        plain_lat.append(clock - a)
    plain_s = time.perf_counter() - t0
    out = {"config": f"{n_req} requests, Poisson {rate:g} jobs/s, 60% Lorenz rho~U[24,32] depth U{{18..26}} k=27, "
                     f"40% forced Duffing depth U{{16..24}} k=16; TM + power iteration to {tol:g}; "
                     f"200 GB admission budget",
           "served": served, "saturated": saturated,
           "plain": {"jobs_per_s": n_req / plain_s, "p50_s": float(np.percentile(plain_lat, 50)),
                     "p99_s": float(np.percentile(plain_lat, 99))},
           "ratio_saturated_vs_plain": saturated["jobs_per_s"] / (n_req / plain_s)}
    if n_steady > 0:
        # a longer stream of the same mix and rate (another seed): the latencies after its first
        # fifth, and the p99 of each fifth, show whether the queue holds steady
        arrival, req = serve_requests(n_steady, rate, seed=1)
        lat, comp, _, info = gx.serve_run(forests(), arrival, *request_args(req))
        k = n_steady // 5
        out["steady"] = {"requests": n_steady, "arrival_span_s": float(arrival[-1]),
                         "jobs_per_s": n_steady / float(comp.max()),
                         "p50_s": float(np.percentile(lat[k:], 50)), "p99_s": float(np.percentile(lat[k:], 99)),
                         "p99_per_fifth_s": [float(np.percentile(lat[i:i + k], 99)) for i in range(0, 5 * k, k)],
                         "max_in_flight": info["max_in_flight"],
                         "note": "latency percentiles over the requests after the first fifth"}
    return out


HH_C, HH_R = [0.0, 0.0, 0.0, 0.0], [0.6, 0.6, 0.5, 0.5]
HH_BAND = (0.11, 0.12)
HH_FLAG = 0x80                   # build-defined flag for the band predicate
HH_FLOPS_PER_IMAGE = 1194        # 50 Verlet steps x 23 + grid/affine/key 44


def hh_loop(t, grid, mp, depth):
    """Config 3 step: subdivide(SD); keep the boxes meeting the energy band; keep the boxes hit
    by an image. Returns the test points imaged."""
    imaged = 0
    t.setFlags(0x08, 0)
    for _ in range(depth):
        t.subdivide(0x08)
        t.setFlagsByEnergyBand(HH_BAND[0], HH_BAND[1], HH_FLAG)
        t.remove(HH_FLAG)
        t.unsetFlags(HH_FLAG, -1)
        imaged += t.size() * grid.points_per_box(4)
        t.setFlagsByIterators(t.makePointIteratorPairOverLeaves(grid, mp), 0x01, -1)
        t.remove(0x01)
        t.unsetFlags(0x01, -1)
    return imaged


HH_VECTORS = 8                   # subspace size: the 4 leading eigenvalues and 4 guard vectors
HH_EIGS_TOL = 1e-10              # Krylov-Schur: relative residual of the leading partial Schur form
HH_EIGS_RESTARTS = 60            # and at most this many restarts


def hh_bench(gx, depth, iters, tol, cpu_depth, threads):
    """Config 3: Hénon-Heiles, 50 Störmer-Verlet steps of h = 0.02, energy band [0.11, 0.12],
    UniformGrid(64, seed 0), the band-restricted subdivision to `depth`, the transition matrix and
    the subspace iteration for the 4 leading eigenvalues (HH_VECTORS vectors), stopped at a
    residual ||X P - H^T X||_F / ||H||_F of `tol` or after `iters` basis updates."""
    import numpy as np
    grid, mp = gx.UniformGrid(64, 0), gx.HenonHeiles(2, 0.02, 50)
    w = gx.ImplicitBoxTree(HH_C, HH_R)  # warm-up pass (kernel loading, image-kernel tuning)
    hh_loop(w, grid, mp, depth)
    w.transitionMatrixForLeaves(grid, mp).subspace_iteration(HH_VECTORS, 1, tol=1.0)
    del w
    t = gx.ImplicitBoxTree(HH_C, HH_R)
    t.image_timing(1)
    t0 = time.perf_counter()
    imaged = hh_loop(t, grid, mp, depth)
    t1 = time.perf_counter()
    m = t.transitionMatrixForLeaves(grid, mp)
    t2 = time.perf_counter()
    H, _, updates, resid = m.subspace_iteration(HH_VECTORS, iters, tol=tol)
    t3 = time.perf_counter()
    # the 4 leading eigenvalues by Krylov-Schur (eigs, 63 vectors), at most HH_EIGS_RESTARTS
    # restarts: the residual reached is reported
    eg = gx.eigs(m, nev=4, tol=HH_EIGS_TOL, max_restarts=HH_EIGS_RESTARTS)
    t4 = time.perf_counter()
    ms_img, launches, pts = t.image_timing(0)
    imaged += t.size() * grid.points_per_box(4)
    ev = np.linalg.eigvals(H)
    ev = ev[np.argsort(-np.abs(ev))]
    out = {"config": f"Hénon-Heiles, 50 Verlet steps h=0.02, band E in [0.11, 0.12], UniformGrid(64, 0), depth {depth}, "
                     f"TM + subspace iteration ({HH_VECTORS} vectors, the 4 leading Ritz values) to residual "
                     f"{tol:g} or {iters} updates",
           "boxes": t.size(), "nnz": m.nnz, "images": imaged, "subspace_updates": updates,
           "subspace_residual": resid, "subspace_converged": bool(resid <= tol),
           "subdiv_s": t1 - t0, "tm_s": t2 - t1, "subspace_s": t3 - t2,
           "images_per_s": imaged / (t3 - t0),
           "images_per_s_subdivision_and_matrix": imaged / (t2 - t0),
           "image_kernel_tflops": pts * HH_FLOPS_PER_IMAGE / (ms_img * 1e-3) / 1e12 if ms_img > 0 else None,
           "ritz_values": [[float(x.real), float(x.imag)] for x in ev[:4]],
           "eigs": {"method": f"Krylov-Schur, {eg['ncv']} basis vectors (b12x_krylov_* on the GPU, eigs() driver)",
                    "tol": HH_EIGS_TOL, "max_restarts": HH_EIGS_RESTARTS, "converged": eg["converged"],
                    "restarts": eg["restarts"],
                    "products": eg["matvecs"], "s": t4 - t3,
                    "eigenvalues": [[float(x.real), float(x.imag)] for x in eg["eigenvalues"]],
                    "residual": eg["residual"], "pair_residuals": [float(x) for x in eg["pair_residuals"]]}}
    if cpu_depth > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_api import OracleTree, set_threads
        set_threads(threads)
        o = OracleTree(HH_C, HH_R)
        o.set_options(eval_twice=True)
        c0 = time.perf_counter()
        ci = hh_loop(o, grid, mp, cpu_depth)
        c1 = time.perf_counter()
        out["cpu_baseline"] = {"value": ci / (c1 - c0), "unit": "images/s", "cores": threads, "kind": "port",
                               "sample": f"the same band-restricted subdivision to depth {cpu_depth} ({ci} images, "
                                         f"{c1 - c0:.1f} s), oracle with the reference's double image evaluation"}
    return out


MANIFOLD_C, MANIFOLD_R = [0.0, 0.0, 27.0], [30.0, 30.0, 40.0]
MANIFOLD_START = [[0.01, 0.01, 0.0], [-0.01, -0.01, 0.0]]  # beside the origin, on either side


def continuation_bench(gx, depth, cpu_depth, threads):
    """Continuation (continuation.m): the Lorenz system's unstable manifold of the origin at one
    depth, InnerGrid(3), started from two points beside the origin. Every box of the final cover
    is mapped exactly once (in the step after its insertion), so the run images final boxes x 27
    test points. Timed through the Python driver over the C ABI, as a caller runs it."""
    import numpy as np
    grid, mp = gx.InnerGrid(3), gx.Lorenz()
    p = np.array(MANIFOLD_START)
    g = gx.ImplicitBoxTree(MANIFOLD_C, MANIFOLD_R)
    gx.continuation(g, grid, mp, [depth], p)  # warm-up (allocations, tuning)
    runs = []
    for _ in range(3):  # the median of three runs
        t0 = time.perf_counter()
        c = gx.continuation(g, grid, mp, [depth], p)[0]
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    out = {"config": f"Lorenz RK4 20x0.01, unstable manifold of the origin, depth {depth}, InnerGrid(3), "
                     f"domain centre {MANIFOLD_C} radius {MANIFOLD_R}",
           "steps": len(c) - 1, "boxes": c[-1], "ms": dt * 1e3, "ms_runs": [x * 1e3 for x in runs],
           "boxes_per_s": c[-1] / dt, "images_per_s": c[-1] * K / dt}
    if cpu_depth > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_api import OracleTree, set_threads
        set_threads(threads)
        o = OracleTree(MANIFOLD_C, MANIFOLD_R)
        c0 = time.perf_counter()
        co = gx.continuation(o, grid, mp, [cpu_depth], p)[0]
        c1 = time.perf_counter()
        out["cpu_baseline"] = {"value": co[-1] * K / (c1 - c0), "unit": "images/s", "cores": threads, "kind": "port",
                               "sample": f"the same continuation at depth {cpu_depth} ({co[-1]} boxes, {len(co) - 1} "
                                         f"steps, {c1 - c0:.1f} s), oracle"}
    return out


def chain_recurrent_bench(gx, steps, cpu, threads):
    """Chain recurrent set (chainRecurrentSet.m) of the Lorenz system: `steps` x {subdivide; the
    leaves' transition matrix (InnerGrid(3)); strong components; keep the leaves of components
    with more than one leaf}, fused on the device. Images: 27 per leaf after each subdivision.
    The strong components of the last cover's matrix are timed on their own as well."""
    grid, mp = gx.InnerGrid(3), gx.Lorenz()

    def run():
        g = gx.ImplicitBoxTree(MANIFOLD_C, MANIFOLD_R)
        g.setFlags(gx.SD, gx.LEAF)
        t0 = time.perf_counter()
        cnt = g.chain_recurrent_steps(grid, mp, steps)
        return g, cnt, time.perf_counter() - t0

    run()  # warm-up
    g, cnt, secs = run()
    leaves = int(cnt[:, 0].sum())
    m = g.transitionMatrixForLeaves(grid, mp)
    m.scc()
    s0 = time.perf_counter()
    _, n_comp, n_big = m.scc()
    s1 = time.perf_counter()
    rounds, passes, checks = gx.api.scc_last_stats()
    out = {"config": f"Lorenz RK4 20x0.01, {steps} steps, InnerGrid(3), domain centre {MANIFOLD_C} radius {MANIFOLD_R}",
           "boxes": int(cnt[-1, 3]), "leaves_stepped": leaves, "ms": secs * 1e3,
           "images_per_s": leaves * K / secs, "boxes_per_s": leaves / secs,
           "scc": {"vertices": m.nRows, "edges": m.nnz, "components": n_comp, "nontrivial": n_big,
                   "ms": (s1 - s0) * 1e3, "rounds": rounds, "passes": passes, "host_checks": checks}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_api import OracleTree, set_threads
        set_threads(threads)
        o = OracleTree(MANIFOLD_C, MANIFOLD_R)
        c0 = time.perf_counter()
        r = gx.chainRecurrentSet(o, grid, mp, steps)
        c1 = time.perf_counter()
        lo = sum(x[0] for x in r)
        out["cpu_baseline"] = {"value": lo * K / (c1 - c0), "unit": "images/s", "cores": threads, "kind": "port",
                               "sample": f"the same {steps} steps ({lo} leaves stepped, {c1 - c0:.1f} s), oracle "
                                         f"with Tarjan's strong components"}
    return out


def host_cores():
    """The cores this process may use (OMP_NUM_THREADS when the machine sets the process's share
    of the host there, else the CPU affinity), the machine's logical CPUs and the CPU model."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    cores = int(env) if env.isdigit() and int(env) > 0 else aff
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(cores, aff), os.cpu_count(), model


def cpu_baseline(sample_depth, threads):
    """B12's CPU path restated (oracle): same algorithm structure as the reference's host
    Thrust path, including its second image evaluation per search, OpenMP over `threads` (B12's
    OMP backend; one thread is its CPP backend)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import b12_amd as gx
    from oracle_api import OracleTree, set_threads
    set_threads(threads)

    o = OracleTree(CENTER, RADIUS)
    o.set_options(eval_twice=True)
    grid, mp = gx.InnerGrid(3), gx.Lorenz()
    o.setFlags(gx.SD, 0)
    t0 = time.perf_counter()
    counts = o.subdivision_steps(grid, mp, sample_depth)
    m = o.transitionMatrixForLeaves(grid, mp)
    _, iters, _ = m.power_iteration(MAX_ITER, TOL)
    t1 = time.perf_counter()
    imgs, n = images_of(counts)
    imgs += K * n
    return imgs / (t1 - t0), t1 - t0, imgs, iters


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--flavour", default="parity", choices=["parity", "fast"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-depth", type=int, default=24)
    ap.add_argument("--cpu1-depth", type=int, default=22, help="depth of the one-thread CPU baseline sample")
    ap.add_argument("--sweep-jobs", type=int, default=512, help="config 4 jobs (0 to skip)")
    ap.add_argument("--serve-requests", type=int, default=200, help="config 5 requests (0 to skip)")
    ap.add_argument("--serve-steady", type=int, default=1000,
                    help="requests of the longer config-5 stream for steady-state latency (0 to skip)")
    ap.add_argument("--hh-depth", type=int, default=28, help="config 3 depth (0 to skip)")
    ap.add_argument("--hh-iters", type=int, default=400, help="config 3 subspace iterations at most")
    ap.add_argument("--hh-tol", type=float, default=1e-8, help="config 3 subspace residual to stop at")
    ap.add_argument("--hh-cpu-depth", type=int, default=18, help="config 3 CPU-baseline sample depth (0 to skip)")
    ap.add_argument("--cont-depth", type=int, default=27, help="continuation depth (0 to skip)")
    ap.add_argument("--cont-cpu-depth", type=int, default=24, help="continuation CPU-baseline depth (0 to skip)")
    ap.add_argument("--chain-steps", type=int, default=30, help="chain recurrent set steps (0 to skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only time a plain device copy of BYTES (traffic-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build's hot path runs on a single MI355X "
              "(one GPU per process; sharding across GPUs is not part of it); run with --gpus 1", file=sys.stderr)
        return 2

    import b12_amd as gx

    if args.copy_calibration:
        ms = gx.bench_copy(args.copy_calibration, 20)
        print(json.dumps({"copy_bytes": args.copy_calibration, "launches": 21, "ms_per_copy": ms,
                          "GB_per_s": 2 * args.copy_calibration / (ms * 1e-3) / 1e9}), flush=True)
        return 0
    dev = gx.device_name()
    for _ in range(args.warmup):
        gpu_job(gx, args.flavour)
    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runs.append(gpu_job(gx, args.flavour))
    t1 = time.perf_counter()
    # image-kernel timing: eager launches bracketed by events on the tree's stream; the first
    # such job settles the autotuner's choices, the second is the one reported
    gpu_job(gx, args.flavour, timing=True)
    prof = gpu_job(gx, args.flavour, timing=True)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    images = runs[0]["images"]
    value = images * args.steps / (t1 - t0)

    ms_img, launches, pts = prof["image_timing"]
    achieved_tflops = pts * LORENZ_FLOPS_PER_IMAGE / (ms_img * 1e-3) / 1e12 if ms_img > 0 else 0.0
    # the shader clock under a full FP64 load (below the 2.4 GHz the peaks assume) and the issue
    # rate at that clock: CUs x 4 SIMDs x 16 lanes x clock
    clock_ghz, clock_probe_frac = gx.api.bench_fp64_clock(args.flavour)
    issue_at_clock = CUS * 64 * clock_ghz * 1e9 / 1e12
    # per launch: test points, microseconds and the fraction of the FP64 issue rate
    levels = [[int(p), round(m * 1e3, 2), round(p * LORENZ_FLOPS_PER_IMAGE / (m * 1e-3) / 1e12
                                                / PEAK_FP64_ISSUE_TFLOPS, 4) if m > 0 else None]
              for m, p in prof["image_launches"]]
    out = {
        "metric": "test-point images/s (Lorenz RK4, depth 24)",
        "value": value,
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seedless InnerGrid test points; the tree is generated on the GPU)",
        "config": {"workload": "lorenz_rk4_depth24_tm_power1e-12", "depth": DEPTH, "points_per_box": K,
                   "map": "Lorenz(10,28,8/3) RK4 20x0.01", "domain": {"center": CENTER, "radius": RADIUS},
                   "build": args.flavour, "device": dev},
        # the job's dominant kernel (about three quarters of its time): the persistent power iteration
        "roofline": measure_roofline(runs[-1]),
        # the image kernel, the job's FP64-bound work: 25 launches (24 subdivision levels and the
        # transition-matrix pass), timed with HIP events on the tree's stream
        "roofline_image": {"bound": "fp64-valu", "kernel": "imageKernel<3,LORENZ,INNER>", "achieved": achieved_tflops,
                           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                           "frac_issue_rate": achieved_tflops / PEAK_FP64_ISSUE_TFLOPS,
                           "clock_ghz_fp64_load": clock_ghz,
                           "frac_issue_rate_at_clock": achieved_tflops / issue_at_clock if issue_at_clock else None,
                           "clock_note": "shader clock measured by b12x_bench_fp64_clock (independent FP64 chains "
                                         f"on every SIMD, which use {clock_probe_frac:.3f} of the issue slots at that "
                                         "clock); the peaks assume 2.4 GHz",
                           "flops_per_image": LORENZ_FLOPS_PER_IMAGE, "image_launches": launches,
                           "image_points": pts, "image_ms": ms_img, "traffic": profiled_traffic(),
                           "traffic_source": "profiles/lorenz_d24_job_traffic.md (FETCH_SIZE x2.000 + WRITE_SIZE x1.000)",
                           "levels": {"columns": ["points", "us", "frac_issue_rate"], "rows": levels}},
        "boxes": {"boxes_per_s": runs[-1]["boxes_stepped"] / (runs[-1]["subdiv_ms"] * 1e-3),
                  "bytes_per_box_step": BYTES_PER_BOX_STEP,
                  "hbm_frac": runs[-1]["boxes_stepped"] * BYTES_PER_BOX_STEP / (runs[-1]["subdiv_ms"] * 1e-3)
                  / PEAK_HBM_BYTES,
                  "note": "leaves processed by the 24 subdivision steps / subdivision wall time"},
        "breakdown_ms": {k: runs[-1][k] for k in ("subdiv_ms", "tm_ms", "measure_ms")},
        "job": {k: runs[-1][k] for k in ("images", "boxes", "nnz", "iters", "resid", "level24_points")},
    }
    def progress(msg):
        print(f"[bench] {msg} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    progress(f"headline: {value:.4g} images/s, {ms_per_step:.2f} ms per job")
    if args.sweep_jobs > 0:
        out["sweep"] = sweep_bench(gx, args.sweep_jobs)
        progress("sweep done")
    if args.serve_requests > 0:
        out["serving"] = serve_bench(gx, args.serve_requests, n_steady=args.serve_steady)
        progress("serving done")
    threads, nproc, cpu_model = host_cores()
    if args.hh_depth > 0:
        out["hh"] = hh_bench(gx, args.hh_depth, args.hh_iters, args.hh_tol,
                             0 if args.no_cpu_baseline else args.hh_cpu_depth, threads)
        progress("config 3 done")
    if args.cont_depth > 0:
        out["continuation"] = continuation_bench(gx, args.cont_depth,
                                                 0 if args.no_cpu_baseline else args.cont_cpu_depth, threads)
        progress("continuation done")
    if args.chain_steps > 0:
        out["chain_recurrent"] = chain_recurrent_bench(gx, args.chain_steps, not args.no_cpu_baseline, threads)
        progress("chain recurrent set done")
    if not args.no_cpu_baseline:
        v, secs, imgs, iters = cpu_baseline(args.cpu_depth, threads)
        v1, secs1, imgs1, iters1 = cpu_baseline(args.cpu1_depth, 1)
        out["cpu_baseline"] = {"value": v, "unit": "images/s", "cores": threads, "kind": "port",
                               "sample": f"same Lorenz job to depth {args.cpu_depth} + transition matrix "
                                         f"and {iters} power iterations ({imgs} images, {secs:.1f} s), oracle "
                                         f"with the reference's double image evaluation per search",
                               "cpu_model": cpu_model, "nproc": nproc,
                               "single_thread": {"value": v1, "unit": "images/s", "cores": 1,
                                                 "sample": f"the same job to depth {args.cpu1_depth} ({imgs1} "
                                                           f"images, {iters1} power iterations, {secs1:.1f} s) on "
                                                           f"one thread, as B12's CPP backend runs"}}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
