"""GPU parity at the full sizes of the Hénon (depth 20) and Lorenz (depth 24) configurations, and
the fast build's accuracy statement: the share of test points whose target box changes under FMA
contraction, and the L1 difference of the invariant measure it causes."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu


def full_job(center, radius, grid, mp, depth, flavour="parity"):
    g = gx.ImplicitBoxTree(center, radius, flavour=flavour)
    g.setFlags(gx.SD, 0)
    counts = g.subdivision_steps(grid, mp, depth)
    return g, counts


def test_henon_depth20_config_matches_oracle():
    """20 subdivision steps of Hénon(1.4, 0.3) with InnerGrid(4), both matrix orders, 200 power
    iterations from the uniform measure."""
    print("[config] henon d20", flush=True)
    c, r, grid, mp = [0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(1.4, 0.3, 1)
    g, cg = full_job(c, r, grid, mp, 20)
    o = OracleTree(c, r)
    o.setFlags(gx.SD, 0)
    co = o.subdivision_steps(grid, mp, 20)
    assert np.array_equal(cg, co)
    assert same_tree(g.download(), o.download())
    for row_major in (False, True):
        mg = g.transitionMatrixForLeaves(grid, mp, row_major=row_major)
        mo = o.transitionMatrixForLeaves(grid, mp, row_major=row_major)
        for a, b in zip(mg.download(), mo.download()):
            assert np.array_equal(a, b)
        mug, itg, rg = mg.power_iteration(200)
        muo, ito, ro = mo.power_iteration(200)
        assert itg == ito == 200 and rg == ro
        assert np.array_equal(mug, muo)  # bit for bit (the stated bound is 1e-12 relative)


def test_lorenz_depth24_config_matches_oracle():
    """24 subdivision steps of Lorenz(10, 28, 8/3; RK4 20 x 0.01) with InnerGrid(3), the
    column-major matrix and power iteration until ||mu_{k+1} - mu_k||_1 <= 1e-12 (the persistent
    straight-line form)."""
    print("[config] lorenz d24", flush=True)
    c, r, grid, mp = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz()
    g, cg = full_job(c, r, grid, mp, 24)
    o = OracleTree(c, r)
    o.setFlags(gx.SD, 0)
    co = o.subdivision_steps(grid, mp, 24)
    assert np.array_equal(cg, co)
    assert same_tree(g.download(), o.download())
    mg = g.transitionMatrixForLeaves(grid, mp)
    mo = o.transitionMatrixForLeaves(grid, mp)
    for a, b in zip(mg.download(), mo.download()):
        assert np.array_equal(a, b)
    mug, itg, rg = mg.power_iteration(1_000_000, 1e-12)
    assert gx.power_iteration_form() == "straight"
    muo, ito, ro = mo.power_iteration(1_000_000, 1e-12)
    print(f"  boxes {g.size()} nnz {mg.nnz} iterations gpu {itg} oracle {ito}", flush=True)
    assert itg == ito and rg == ro and rg <= 1e-12
    assert np.array_equal(mug, muo)  # bit for bit (the stated bound is 1e-12 relative)


def test_fast_build_target_changes_and_measure_difference():
    """The fast build (FMA contraction in the integrator) on the parity build's Lorenz depth-24
    cover: at most 1e-6 of the test points may change target box; the L1 difference between the
    two invariant measures (same cover, to 1e-12) is reported."""
    print("[config] fast build", flush=True)
    c, r, grid, mp = [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz()
    g, _ = full_job(c, r, grid, mp, 24)
    bits, depth, flags = g.download()
    f = gx.ImplicitBoxTree(c, r, flavour="fast")
    f.upload(bits, depth, flags)
    tg = g.searchByIterators(g.makePointIteratorPairOverLeaves(grid, mp))
    tf = f.searchByIterators(f.makePointIteratorPairOverLeaves(grid, mp))
    changed = int(np.count_nonzero(tg != tf))
    share = changed / len(tg)
    mug, _, _ = g.transitionMatrixForLeaves(grid, mp).power_iteration(1_000_000, 1e-12)
    muf, _, _ = f.transitionMatrixForLeaves(grid, mp).power_iteration(1_000_000, 1e-12)
    l1 = float(np.abs(mug - muf).sum())
    print(f"  fast build: {changed} of {len(tg)} test points change target box (share {share:.3g}); "
          f"measure L1 difference {l1:.3g}", flush=True)
    assert share <= 1e-6
    assert l1 <= 1e-6


def test_config4_sweep_of_512_jobs_matches_single_trees():
    """Configuration 4 at its size: the 512 Hénon jobs of the benchmark's sweep (a ~ U[1.0, 1.4],
    b = 0.3, depth ~ U{14..22}, InnerGrid(4), 200 power iterations) in one forest. Every job's
    cover, transition matrix and invariant measure equal its single-tree run bit for bit, and the
    deepest jobs equal the oracle's."""
    print("[config] 512-job sweep", flush=True)
    import bench
    from oracle_api import OracleTree as OT
    jobs = bench.sweep_jobs(512)
    grid = gx.InnerGrid(4)
    F = gx.Forest(2, grid, 1, pi_max_iter=200)
    ids = [F.add(bench.HENON_C, bench.HENON_R, gx.Henon(a, 0.3, 1), d) for a, d in jobs]
    F.run()
    sizes = []
    for j, (a, d) in zip(ids, jobs):
        mp = gx.Henon(a, 0.3, 1)
        t = gx.ImplicitBoxTree(bench.HENON_C, bench.HENON_R)
        t.setFlags(gx.SD, 0)
        t.subdivision_steps(grid, mp, d)
        m = t.transitionMatrixForLeaves(grid, mp)
        mu, it, _ = m.power_iteration(200)
        assert F.info(j)["state"] == 2 and F.info(j)["iters"] == it == 200
        assert same_tree(F.cover(j), t.download())
        for x, y in zip(F.matrix(j), m.download()):
            assert np.array_equal(x, y)
        assert np.array_equal(F.measure(j), mu)
        sizes.append(t.size())
    print(f"  covers from {min(sizes)} to {max(sizes)} boxes", flush=True)
    deepest = sorted(range(512), key=lambda q: -jobs[q][1])[:3]
    for q in deepest:
        a, d = jobs[q]
        o = OT(bench.HENON_C, bench.HENON_R)
        o.setFlags(gx.SD, 0)
        o.subdivision_steps(grid, gx.Henon(a, 0.3, 1), d)
        mo = o.transitionMatrixForLeaves(grid, gx.Henon(a, 0.3, 1))
        muo, ito, _ = mo.power_iteration(200)
        assert same_tree(F.cover(ids[q]), o.download())
        for x, y in zip(F.matrix(ids[q]), mo.download()):
            assert np.array_equal(x, y)
        assert np.array_equal(F.measure(ids[q]), muo)
