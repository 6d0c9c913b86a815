"""GPU tests of the batched and served forms: every job's cover, transition matrix and invariant
measure equal the single-tree results bit for bit (and the oracle's, with the same iteration
count), and a job's results are identical alone and in any mix."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu

HENON_C, HENON_R = [0.0, 0.0], [1.5, 0.5]


def oracle_job(center, radius, grid, mp, depth, pi_iter=200, tol=-1.0, row_major=False):
    o = OracleTree(center, radius)
    o.setFlags(gx.SD, 0)
    o.subdivision_steps(grid, mp, depth)
    m = o.transitionMatrixForLeaves(grid, mp, row_major=row_major)
    mu, it, res = m.power_iteration(pi_iter, tol)
    return o.download(), m.download(), mu, it


def assert_job_matches(forest, job, ref):
    cover, mat, mu, it = ref
    assert same_tree(forest.cover(job), cover)
    for a, b in zip(forest.matrix(job), mat):
        assert np.array_equal(a, b)
    got = forest.measure(job)
    assert forest.info(job)["iters"] == it
    assert np.array_equal(got, mu)  # bit for bit (the stated bound is 1e-12 relative)
    return got


def test_batched_henon_sweep_matches_single_jobs():
    print("[forest] henon sweep", flush=True)
    rng = np.random.default_rng(0)
    grid = gx.InnerGrid(4)
    F = gx.Forest(2, grid, 1, pi_max_iter=200)
    jobs = []
    for _ in range(10):
        a = float(rng.uniform(1.0, 1.4))
        d = int(rng.integers(6, 13))
        jobs.append((F.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d), a, d))
    F.run()
    for j, a, d in jobs:
        info = F.info(j)
        assert info["state"] == 2 and info["iters"] == 200
        assert_job_matches(F, j, oracle_job(HENON_C, HENON_R, grid, gx.Henon(a, 0.3, 1), d))


def test_job_identical_alone_and_in_the_mix():
    grid = gx.InnerGrid(4)
    mp = gx.Henon(1.3, 0.3, 1)
    alone = gx.Forest(2, grid, 1, pi_max_iter=150)
    ja = alone.add(HENON_C, HENON_R, mp, 11)
    alone.run()
    mixed = gx.Forest(2, grid, 1, pi_max_iter=150)
    rng = np.random.default_rng(1)
    for _ in range(5):
        mixed.add(HENON_C, HENON_R, gx.Henon(float(rng.uniform(1.0, 1.4)), 0.3, 1), int(rng.integers(4, 14)))
    jm = mixed.add(HENON_C, HENON_R, mp, 11)
    for _ in range(3):
        mixed.step()
    for _ in range(4):  # jobs joining while others are in flight
        mixed.add(HENON_C, HENON_R, gx.Henon(float(rng.uniform(1.0, 1.4)), 0.3, 1), int(rng.integers(3, 9)))
    mixed.run()
    assert same_tree(alone.cover(ja), mixed.cover(jm))
    for a, b in zip(alone.matrix(ja), mixed.matrix(jm)):
        assert np.array_equal(a, b)
    assert np.array_equal(alone.measure(ja), mixed.measure(jm))
    assert alone.info(ja)["iters"] == mixed.info(jm)["iters"] == 150


def test_lorenz_forest_tolerance_and_row_major():
    print("[forest] lorenz", flush=True)
    grid = gx.InnerGrid(3)
    c, r = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0]
    for row_major in (False, True):
        F = gx.Forest(3, grid, 2, row_major=row_major, pi_max_iter=100000, pi_tol=1e-12)
        jobs = [(F.add(c, r, gx.Lorenz(rho=rho), d), rho, d) for rho, d in [(28.0, 10), (24.5, 9), (31.0, 11)]]
        F.run()
        for j, rho, d in jobs:
            ref = oracle_job(c, r, grid, gx.Lorenz(rho=rho), d, 100000, 1e-12, row_major)
            cover, mat, mu, it = ref
            assert same_tree(F.cover(j), cover)
            for a, b in zip(F.matrix(j), mat):
                assert np.array_equal(a, b)
            info = F.info(j)
            assert info["iters"] == it and info["resid"] <= 1e-12
            assert np.array_equal(F.measure(j), mu)


def test_duffing_and_uniform_grid_forests():
    grid = gx.InnerGrid(4)
    F = gx.Forest(2, grid, 4, pi_max_iter=50)
    j = F.add([0.0, 0.0], [2.0, 2.0], gx.Duffing(), 8)
    F.run()
    assert_job_matches(F, j, oracle_job([0.0, 0.0], [2.0, 2.0], grid, gx.Duffing(), 8, 50))
    ug = gx.UniformGrid(16, 7)
    G = gx.Forest(2, ug, 1, pi_max_iter=30)
    jobs = [G.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d) for a, d in [(1.4, 7), (1.2, 8)]]
    G.run()
    for jj, (a, d) in zip(jobs, [(1.4, 7), (1.2, 8)]):
        assert_job_matches(G, jj, oracle_job(HENON_C, HENON_R, ug, gx.Henon(a, 0.3, 1), d, 30))


def test_serve_run_results_match_alone_runs():
    print("[serve] small stream", flush=True)
    rng = np.random.default_rng(2)
    lor = gx.Forest(3, gx.InnerGrid(3), 2, pi_max_iter=100000, pi_tol=1e-12)
    duf = gx.Forest(2, gx.InnerGrid(4), 4, pi_max_iter=100000, pi_tol=1e-12)
    n = 12
    arrival = np.cumsum(rng.exponential(1 / 200.0, n))
    kinds = (rng.random(n) < 0.6).astype(int)  # 1 = Lorenz
    forest_of, centers, radii, params, depths = [], [], [], [], []
    for k in kinds:
        if k:
            forest_of.append(0); centers.append([0.0, 0.0, 27.0]); radii.append([30.0, 30.0, 30.0])
            params.append(gx.Lorenz(rho=float(rng.uniform(24, 32))).params); depths.append(int(rng.integers(6, 11)))
        else:
            forest_of.append(1); centers.append([0.0, 0.0]); radii.append([2.0, 2.0])
            params.append(gx.Duffing().params); depths.append(int(rng.integers(5, 9)))
    lat, comp, ids, _ = gx.serve_run([lor, duf], arrival, forest_of, centers, radii, params, depths)
    assert np.all(lat > 0) and np.all(comp >= arrival)
    for i in range(n):
        F = lor if forest_of[i] == 0 else duf
        dim = 3 if forest_of[i] == 0 else 2
        kind = 2 if forest_of[i] == 0 else 4
        grid = gx.InnerGrid(3) if dim == 3 else gx.InnerGrid(4)
        alone = gx.Forest(dim, grid, kind, pi_max_iter=100000, pi_tol=1e-12)
        mp = gx.Lorenz(*params[i]) if dim == 3 else gx.Duffing(*params[i])
        ja = alone.add(centers[i], radii[i], mp, depths[i])
        alone.run()
        assert same_tree(alone.cover(ja), F.cover(int(ids[i])))
        for a, b in zip(alone.matrix(ja), F.matrix(int(ids[i]))):
            assert np.array_equal(a, b)
        assert np.array_equal(alone.measure(ja), F.measure(int(ids[i])))


def test_pool_graph_replays_match_eager_launches(monkeypatch):
    """The pool's captured batches (replayed many times, recaptured as jobs join and leave) give
    the same measures and iteration counts as eager launches."""
    print("[forest] pool graph vs eager", flush=True)
    grid = gx.InnerGrid(3)
    c, r = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0]
    cases = [(28.0, 10), (24.5, 9), (31.0, 11), (26.0, 8), (29.5, 12)]
    out = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("B12X_GRAPHS", graphs)
        F = gx.Forest(3, grid, 2, pi_max_iter=100000, pi_tol=1e-12)
        F.set_pool_iterations(16)
        jobs = [F.add(c, r, gx.Lorenz(rho=rho), d) for rho, d in cases[:3]]
        F.step()
        jobs += [F.add(c, r, gx.Lorenz(rho=rho), d) for rho, d in cases[3:]]
        F.run()
        out.append([(F.measure(j), F.info(j)["iters"]) for j in jobs])
        assert F.stats()["pool_iters"] > 16
    for (ma, ia), (mb, ib) in zip(*out):
        assert ia == ib
        assert np.array_equal(ma, mb)


def test_persistent_and_batched_pool_forms_agree(monkeypatch):
    """A pool larger than the persistent form takes (batched launches) that shrinks below it as
    jobs leave (persistent launches): every job's measure and iteration count equal the same
    job run alone (persistent throughout), the all-batched run and a run whose persistent waits
    give up at once (the jobs still running restart from iteration 1 in the batched form), bit
    for bit."""
    print("[forest] persistent vs batched pool", flush=True)
    grid = gx.InnerGrid(4)
    rng = np.random.default_rng(9)
    cases = [(float(rng.uniform(1.2, 1.4)), int(rng.integers(20, 22))) for _ in range(14)]
    names = ("B12X_POOL_PERSISTENT", "B12X_POOL_SPIN")

    def run(**env):
        for k in names:
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        F = gx.Forest(2, grid, 1, pi_max_iter=100000, pi_tol=1e-12)
        F.set_pool_iterations(8)
        ids = [F.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d) for a, d in cases]
        F.run()
        st = F.stats()
        print(f"  {env}: pool iterations {st['pool_iters']:.0f}", flush=True)
        return [(F.measure(j), F.info(j)["iters"]) for j in ids], st

    mixed, st = run()
    assert st["pool_max_cols"] > 512 * 256  # the pool started above the persistent form's size
    batched, _ = run(B12X_POOL_PERSISTENT="0")
    gave_up, _ = run(B12X_POOL_SPIN="0")
    for other in (batched, gave_up):
        for (ma, ia), (mb, ib) in zip(mixed, other):
            assert ia == ib and np.array_equal(ma, mb)
    for k in names:
        monkeypatch.delenv(k, raising=False)
    a, d = cases[0]
    A = gx.Forest(2, grid, 1, pi_max_iter=100000, pi_tol=1e-12)
    A.set_pool_iterations(8)
    j = A.add(HENON_C, HENON_R, gx.Henon(a, 0.3, 1), d)
    A.run()
    assert A.info(j)["iters"] == mixed[0][1] and np.array_equal(A.measure(j), mixed[0][0])
    ref = oracle_job(HENON_C, HENON_R, grid, gx.Henon(a, 0.3, 1), d, 100000, 1e-12)
    assert A.info(j)["iters"] == ref[3]
    assert np.array_equal(A.measure(j), ref[2])


def test_config5_served_jobs_at_the_configuration_depths():
    """Configuration 5's jobs at its own depths, served from a Poisson stream at 200 jobs/s:
    Lorenz (rho ~ U[24, 32], InnerGrid(3)) at depths 18, 20 and 22, forced Duffing (period map in
    100 RK4 steps, InnerGrid(4)) at depths 16 and 19, measures to 1e-12. Each served result equals
    the same job run alone bit for bit, and the Lorenz jobs equal the oracle's (cover, matrix,
    measure and iteration count)."""
    print("[serve] configuration 5 depths", flush=True)
    rng = np.random.default_rng(5)
    lor_c, lor_r, duf_c, duf_r = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], [0.0, 0.0], [2.0, 2.0]
    reqs = [(0, 18), (1, 16), (0, 20), (1, 19), (0, 22)]
    forest_of, centers, radii, params, depths = [], [], [], [], []
    for q, d in reqs:
        forest_of.append(q)
        centers.append(lor_c if q == 0 else duf_c)
        radii.append(lor_r if q == 0 else duf_r)
        params.append(gx.Lorenz(rho=float(rng.uniform(24, 32))).params if q == 0 else gx.Duffing().params)
        depths.append(d)
    arrival = np.cumsum(rng.exponential(1 / 200.0, len(reqs)))

    def forests():
        return [gx.Forest(3, gx.InnerGrid(3), 2, pi_max_iter=1_000_000, pi_tol=1e-12),
                gx.Forest(2, gx.InnerGrid(4), 4, pi_max_iter=1_000_000, pi_tol=1e-12)]

    fs = forests()
    lat, comp, ids, _ = gx.serve_run(fs, arrival, forest_of, centers, radii, params, depths)
    assert np.all(lat > 0) and np.all(comp >= arrival)
    for i, (q, d) in enumerate(reqs):
        F = fs[q]
        grid = gx.InnerGrid(3) if q == 0 else gx.InnerGrid(4)
        mp = gx.Lorenz(*params[i]) if q == 0 else gx.Duffing(*params[i])
        alone = forests()[q]
        ja = alone.add(centers[i], radii[i], mp, d)
        alone.run()
        j = int(ids[i])
        print(f"  request {i}: {'Lorenz' if q == 0 else 'Duffing'} depth {d}, {F.info(j)['n']} boxes, "
              f"{F.info(j)['iters']} iterations", flush=True)
        assert same_tree(alone.cover(ja), F.cover(j))
        for a, b in zip(alone.matrix(ja), F.matrix(j)):
            assert np.array_equal(a, b)
        assert alone.info(ja)["iters"] == F.info(j)["iters"]
        assert np.array_equal(alone.measure(ja), F.measure(j))
        if q == 0:
            cover, mat, mu, it = oracle_job(centers[i], radii[i], grid, mp, d, 1_000_000, 1e-12)
            assert same_tree(F.cover(j), cover)
            for a, b in zip(F.matrix(j), mat):
                assert np.array_equal(a, b)
            assert F.info(j)["iters"] == it and np.array_equal(F.measure(j), mu)


def test_serve_admission_under_a_binding_budget():
    """Admission by projected HBM: a budget of the forests' base footprint plus 3.2 times one
    job's projected cost (leaves x (208 + 4k) + entries x 74 bytes, the rule b12x_serve_run
    documents) admits at most 3 of 9 simultaneous Lorenz requests at once; the others wait and
    are admitted as earlier jobs complete; the forests' footprint never exceeds the budget; every
    served result equals the same job run alone, bit for bit."""
    print("[serve] binding budget", flush=True)
    c, r, grid, depth = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), 14
    rhos = [24.0 + 0.9 * q for q in range(9)]

    def forest():
        return gx.Forest(3, grid, 2, pi_max_iter=1_000_000, pi_tol=1e-12)

    alone = []
    cost = 0.0
    for rho in rhos:
        A = forest()
        ja = A.add(c, r, gx.Lorenz(rho=rho), depth)
        A.run()
        info = A.info(ja)
        cost = max(cost, info["n"] * (208 + 4 * 27) + info["nnz"] * 74.0)
        alone.append((A, ja))
    F = forest()
    base = F.footprint()
    budget = base + 3.2 * cost
    n = len(rhos)
    lat, comp, ids, info = gx.serve_run([F], np.zeros(n), [0] * n, [c] * n, [r] * n,
                                        [gx.Lorenz(rho=rho).params for rho in rhos], [depth] * n,
                                        budget_bytes=budget, reserve_per_job=cost)
    adm = info["admitted"]
    print(f"  cost {cost / 1e6:.2f} MB budget {budget / 1e6:.2f} MB peak {info['peak_footprint'] / 1e6:.2f} MB "
          f"max in flight {info['max_in_flight']} deferred {info['deferred']}", flush=True)
    assert 1 <= info["max_in_flight"] <= 3 and info["deferred"] >= 1
    assert info["peak_footprint"] <= budget
    # at every admission after the first three, an earlier request had completed
    order = np.argsort(adm, kind="stable")
    for q in order[3:]:
        assert np.any((comp <= adm[q]) & (np.arange(n) != q))
    for q, (A, ja) in enumerate(alone):
        j = int(ids[q])
        assert same_tree(A.cover(ja), F.cover(j))
        for a, b in zip(A.matrix(ja), F.matrix(j)):
            assert np.array_equal(a, b)
        assert A.info(ja)["iters"] == F.info(j)["iters"] and np.array_equal(A.measure(ja), F.measure(j))
