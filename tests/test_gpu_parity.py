"""GPU parity tests: the HIP library (through its C ABI) against the CPU oracle on the same
seeded inputs. Integer, key and index results must match bit for bit; FP64 map images must be
bit-identical (parity build, contraction off); the invariant measure within 1e-12 relative."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree
from oracle_api import OracleTree, OCoo

pytestmark = pytest.mark.gpu

MEASURE_RTOL = 1e-12


def both(center, radius, sd=None):
    return gx.ImplicitBoxTree(center, radius, sd), OracleTree(center, radius, sd)


def assert_same_state(g, o):
    a, b = g.download(), o.download()
    assert same_tree(a, b), "leaf arrays differ"


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name,dim,mp,scale,shift", [
    ("lorenz", 3, gx.Lorenz(), 20.0, (0, 0, 27)),
    ("lorenz_rho", 3, gx.Lorenz(rho=24.5, steps=7), 20.0, (0, 0, 27)),
    ("henon1", 2, gx.Henon(1.4, 0.3, 1), 1.5, (0, 0)),
    ("henon4", 2, gx.Henon(1.2, 0.3, 4), 1.0, (0, 0)),
    ("henon3", 2, gx.Henon(1.0, 0.3, 3), 1.0, (0, 0)),
    ("hh2", 4, gx.HenonHeiles(2, 0.02, 50), 0.4, (0, 0, 0, 0)),
    ("hh1", 4, gx.HenonHeiles(1, 0.02, 10), 0.4, (0, 0, 0, 0)),
    ("hh3", 4, gx.HenonHeiles(3, 0.02, 10), 0.4, (0, 0, 0, 0)),
    ("hh4", 4, gx.HenonHeiles(4, 0.02, 10), 0.4, (0, 0, 0, 0)),
    ("duffing", 2, gx.Duffing(), 1.5, (0, 0)),
    ("identity3", 3, gx.Identity(3), 5.0, (0, 0, 0)),
])
def test_map_images_bit_exact(name, dim, mp, scale, shift):
    print(f"[map] {name}", flush=True)
    rng = np.random.default_rng(1)
    x = rng.uniform(-scale, scale, (4096, dim)) + np.asarray(shift, float)
    g = gx.Scheme(dim, gx.InnerGrid(2), mp)
    o = OracleTree(np.zeros(dim), np.ones(dim)).scheme(gx.InnerGrid(2), mp)
    yg, yo = g.map_points(x), o.map_points(x)
    assert np.array_equal(yg.view(np.uint64), yo.view(np.uint64))


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim,seed", [(1, 0), (2, 1), (3, 2), (4, 3), (2, 4), (3, 5)])
def test_tree_operations_random(dim, seed):
    print(f"[ops] dim={dim} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, dim)
    r = rng.uniform(0.5, 2, dim)
    g, o = both(c, r)
    bits, depth, flags = random_tree(rng, dim, 300, max_depth=10)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    assert_same_state(g, o)
    ops = [
        ("subdivide", (8,)), ("setFlags", (1, -1)), ("unsetFlags", (8, -1)), ("changeFlags", (1, 0x10, -1)),
        ("subdivide", (0x10,)), ("remove", (0x10,)), ("unsubdivide", (0x10,)), ("setFlags", (8, 0)),
        ("unsubdivide", (8,)), ("unsetFlags", (0x10, 0)), ("changeFlags", (8, 1, 0)), ("subdivide", (1,)),
        ("remove", (0x40,)), ("subdivide", (0,)), ("unsubdivide", (0,)),
    ]
    for name, args in ops:
        rg, ro = getattr(g, name)(*args), getattr(o, name)(*args)
        assert rg == ro, f"{name}{args}: {rg} != {ro}"
        assert_same_state(g, o)
        assert g.count(-1) == o.count(-1) and g.depth() == o.depth()
    n = g.size()
    st = rng.random(n) < 0.5
    for name, args in [("setFlagsByStencil", (st, 0x20)), ("unsetFlagsByStencil", (st, 0x20)),
                       ("changeFlagsByStencil", (st, 0x08, 0x40))]:
        assert getattr(g, name)(*args) == getattr(o, name)(*args)
        assert_same_state(g, o)


def test_search_by_points_and_invalid_quirk():
    rng = np.random.default_rng(7)
    g, o = both([0.5, -0.25], [1.5, 0.75])
    bits, depth, flags = random_tree(rng, 2, 2000, max_depth=16)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    pts = rng.uniform(-1.5, 2.5, (20000, 2))
    pts[:50] = [0.5 - 1.5, -1.0]  # exactly on the lower domain edge
    assert np.array_equal(g.searchByPoints(pts), o.searchByPoints(pts))
    assert np.array_equal(g.searchByPoints(pts, 0), o.searchByPoints(pts, 0))
    # single all-zero leaf: points outside the domain are reported on leaf 0 (B12 quirk)
    g2, o2 = both([0.0, 0.0], [1.0, 1.0])
    for t in (g2, o2):
        t.upload(np.array([0], np.uint64), np.array([3], np.uint8), np.array([0], np.uint8))
    p = np.array([[5.0, 5.0], [-0.9, -0.9], [0.5, 0.5]])
    rg, ro = g2.searchByPoints(p), o2.searchByPoints(p)
    assert np.array_equal(rg, ro) and rg[0] == 0 and rg[2] == gx.INVALID


@pytest.mark.parametrize("n", [1, 3, 2047, 2048, 2049, 6000, 30000])
def test_descent_staged_in_lds_matches_global_descent(monkeypatch, n):
    """The descent with its upper levels staged in LDS (samples of the leaf keys) gives the same
    leaf indices as the all-global binary search and as the oracle, on trees with holes of sizes
    around the sample count, with points outside the domain and on the domain edge."""
    print(f"[descent] n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    bits, depth, flags = random_tree(rng, 2, int(n * 1.3) + 2, max_depth=22)
    pick = rng.permutation(len(bits))[:n]
    bits, depth, flags = bits[pick], depth[pick], flags[pick]
    g, o = both([0.0, 0.0], [1.5, 1.5])
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    assert g.size() == n
    mp = gx.Henon(1.4, 0.3, 1)
    ig = g.makePointIteratorPairOverLeaves(gx.InnerGrid(4), mp)
    io = o.makePointIteratorPairOverLeaves(gx.InnerGrid(4), mp)
    ref = o.searchByIterators(io)
    plain = g.searchByIterators(ig)
    monkeypatch.setenv("B12X_DESCENT_LDS", "1")
    staged = g.searchByIterators(ig)
    assert np.array_equal(g.searchCompacted(ig), o.searchCompacted(io))
    monkeypatch.delenv("B12X_DESCENT_LDS")
    assert np.array_equal(staged, ref) and np.array_equal(plain, ref)


# ---------------------------------------------------------------------------------------------
def subdivision_run(center, radius, grid, mp, steps):
    g, o = both(center, radius)
    assert g.setFlags(gx.SD, 0) == o.setFlags(gx.SD, 0)
    for s in range(steps):
        cg = g.subdivision_steps(grid, mp, 1)[0]
        co = o.subdivision_steps(grid, mp, 1)[0]
        assert np.array_equal(cg, co), f"step {s}: counts {cg} vs {co}"
        assert_same_state(g, o)
    return g, o


def test_subdivision_henon_cover():
    print("[subdivision] henon d14", flush=True)
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 14)
    assert g.count() > 500


def test_subdivision_lorenz_cover():
    print("[subdivision] lorenz d12", flush=True)
    g, o = subdivision_run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 12)
    assert g.count() > 100


def test_subdivision_duffing_cover():
    print("[subdivision] duffing d10", flush=True)
    subdivision_run([0.0, 0.0], [2.0, 2.0], gx.InnerGrid(4), gx.Duffing(), 10)


@pytest.mark.parametrize("name,center,radius,grid,mp,steps", [
    ("henon", [0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 18),
    ("lorenz", [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 26),
])
def test_subdivision_step_graph_matches_eager_launches(monkeypatch, name, center, radius, grid, mp, steps):
    """The batched subdivision steps (levels whose subdivided count fits a batch form's bound),
    launched on the stream or replayed from a captured graph, and the eager launches with a host
    synchronisation per step give the same per-step counts and covers; the Lorenz run crosses
    the bound (more than 65536 leaves from depth 25 on), so the forms alternate within it."""
    print(f"[subdivision batches] {name} d{steps}", flush=True)
    res = {}
    for mode, graphs, launch in (("stream", "1", "stream"), ("graph", "1", "graph"), ("eager", "0", "stream")):
        monkeypatch.setenv("B12X_GRAPHS", graphs)
        monkeypatch.setenv("B12X_STEP_LAUNCH", launch)
        g = gx.ImplicitBoxTree(center, radius)
        g.setFlags(gx.SD, 0)
        res[mode] = (g.subdivision_steps(grid, mp, steps), g.download())
    for mode in ("graph", "eager"):
        assert np.array_equal(res["stream"][0], res[mode][0]), mode
        assert same_tree(res["stream"][1], res[mode][1]), mode
    if name == "lorenz":
        assert res["stream"][0][-1][3] > 65536


@pytest.mark.parametrize("sd_flag", [gx.SD, 0])
def test_subdivision_steps_that_empty_the_tree(monkeypatch, sd_flag):
    """Steps whose images all leave the domain keep no leaf, so the surrounding box comes back as
    the only leaf (ImplicitBoxTree.hpp:1959-1970) and the next step starts from it. With the SD
    flag the box, which carries no flags, stays whole, and the images outside the domain land
    on it (a tree whose only leaf is the surrounding box reports them there), so it is kept;
    with no flag (0) every leaf, the box included, is subdivided again and nothing is kept. The
    batched steps (several in one batch, on the stream or from a captured graph), the eager
    launches and the oracle agree on every step's counts and on the leaves."""
    print(f"[subdivision] emptied tree, sd flag {sd_flag}", flush=True)
    center, radius = [10.0, 10.0], [1.0, 1.0]  # the Hénon map sends the whole box outside it
    grid, mp, steps = gx.InnerGrid(4), gx.Henon(), 6
    res = {}
    for mode, graphs, launch in (("stream", "1", "stream"), ("graph", "1", "graph"), ("eager", "0", "stream")):
        monkeypatch.setenv("B12X_GRAPHS", graphs)
        monkeypatch.setenv("B12X_STEP_LAUNCH", launch)
        g = gx.ImplicitBoxTree(center, radius)
        g.setFlags(gx.SD, 0)
        res[mode] = (g.subdivision_steps(grid, mp, steps, sd_flag), g.download())
    o = OracleTree(center, radius)
    o.setFlags(gx.SD, 0)
    co = o.subdivision_steps(grid, mp, steps, sd_flag)
    assert co[0, 3] == 0  # the first step keeps no leaf
    assert co[0, 0] == 1 and (co[1:, 0] == (1 if sd_flag == 0 else 0)).all()
    if sd_flag == 0:
        assert (co[:, 3] == 0).all()
    ref = o.download()
    for mode in res:
        assert np.array_equal(res[mode][0], co), (mode, res[mode][0], co)
        assert same_tree(res[mode][1], ref), mode


def test_stepwise_api_matches_fused_step():
    """The four reference calls one by one give the same cover as the fused step."""
    g1, o = both([0.0, 0.0], [1.5, 0.5])
    g2 = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    for t in (g1, g2, o):
        t.setFlags(gx.SD, 0)
    grid, mp = gx.InnerGrid(4), gx.Henon()
    for _ in range(10):
        r1 = [g1.subdivide(gx.SD)]
        it = g1.makePointIteratorPairOverLeaves(grid, mp)
        r1 += [g1.setFlagsByIterators(it, gx.HIT, -1), g1.remove(gx.HIT), g1.unsetFlags(gx.HIT, -1)]
        r2 = g2.subdivision_steps(grid, mp, 1)[0]
        r3 = o.subdivision_steps(grid, mp, 1)[0]
        assert list(r1) == list(r2) == list(r3)
        assert same_tree(g1.download(), o.download()) and same_tree(g2.download(), o.download())


def test_search_iterators_lorenz():
    g, o = subdivision_run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 8)
    for mp in (gx.Lorenz(), gx.Lorenz(rho=26.0, steps=10)):
        ig = g.makePointIteratorPairOverLeaves(gx.InnerGrid(3), mp)
        io = o.makePointIteratorPairOverLeaves(gx.InnerGrid(3), mp)
        assert np.array_equal(g.searchByIterators(ig), o.searchByIterators(io))
        assert np.array_equal(g.searchCompacted(ig), o.searchCompacted(io))
        rg = g.makePointIteratorPairOverRoot(gx.InnerGrid(3), mp)
        ro = o.makePointIteratorPairOverRoot(gx.InnerGrid(3), mp)
        assert np.array_equal(g.searchByIterators(rg), o.searchByIterators(ro))
        assert np.array_equal(g.searchByIterators(ig, 0), o.searchByIterators(io, 0))


def test_uniform_grid_search_hh():
    rng = np.random.default_rng(3)
    g, o = both([0.0, 0.0, 0.0, 0.0], [0.6, 0.6, 0.5, 0.5])
    bits, depth, flags = random_tree(rng, 4, 1500, max_depth=14)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    for grid in (gx.UniformGrid(64, 0), gx.UniformGrid(5, 12345)):
        ig = g.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
        io = o.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
        assert np.array_equal(g.searchByIterators(ig), o.searchByIterators(io))


def test_by_flags_iterator_and_flag_ops():
    rng = np.random.default_rng(11)
    g, o = both([0.0, 0.0], [1.5, 0.5])
    bits, depth, flags = random_tree(rng, 2, 800, max_depth=12)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    for check in (0x08, 0x09, 0x00):
        ig = g.makePointIteratorPairByFlags(-1, check, gx.InnerGrid(3), gx.Henon())
        io = o.makePointIteratorPairByFlags(-1, check, gx.InnerGrid(3), gx.Henon())
        assert g.num_points(ig) == o.num_points(io)
        assert np.array_equal(g.searchByIterators(ig), o.searchByIterators(io))
        assert g.setFlagsByIterators(ig, 0x20) == o.setFlagsByIterators(io, 0x20)
        assert g.changeFlagsByIterators(ig, 0x20, 0x40) == o.changeFlagsByIterators(io, 0x20, 0x40)
        assert g.unsetFlagsByIterators(ig, 0x40) == o.unsetFlagsByIterators(io, 0x40)
        assert_same_state(g, o)
        assert g.setFlagsByIterators(ig, 0x02, 0) == o.setFlagsByIterators(io, 0x02, 0)
        assert_same_state(g, o)


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("use_abs,row_major,use_invalid", [
    (False, False, False), (False, True, False), (True, False, False), (True, True, True), (False, False, True),
    (False, True, True)])
def test_transition_matrix_henon(use_abs, row_major, use_invalid):
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 10)
    # remove half the leaves so that some targets fall outside the cover (invalid)
    st = np.arange(g.size()) % 3 != 0
    g.setFlagsByStencil(st, 0x20)
    o.setFlagsByStencil(st, 0x20)
    g.remove(0x20)
    o.remove(0x20)
    mg = g.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), use_abs, row_major, use_invalid)
    mo = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), use_abs, row_major, use_invalid)
    assert (mg.nnz, mg.nRows, mg.nColumns) == (mo.nnz, mo.nRows, mo.nColumns)
    for a, b in zip(mg.download(), mo.download()):
        assert np.array_equal(a, b)


def test_transition_matrix_lorenz_and_root():
    g, o = subdivision_run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 9)
    mg = g.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz())
    mo = o.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz())
    for a, b in zip(mg.download(), mo.download()):
        assert np.array_equal(a, b)
    rg = g.transitionMatrixForRoot(gx.InnerGrid(3), gx.Lorenz(), use_abs=True)
    ro = o.transitionMatrixForRoot(gx.InnerGrid(3), gx.Lorenz(), use_abs=True)
    for a, b in zip(rg.download(), ro.download()):
        assert np.array_equal(a, b)


def test_power_iteration_matches_oracle():
    """Fixed counts and a 1e-12 tolerance, both matrix orders: the same iteration count, residual
    and measure as the oracle, bit for bit (the stated bound is 1e-12 relative)."""
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 12)
    for rm in (False, True):
        mg = g.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), row_major=rm)
        mo = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), row_major=rm)
        mug, itg, resg = mg.power_iteration(200)
        muo, ito, reso = mo.power_iteration(200)
        assert itg == ito == 200 and resg == reso
        assert np.array_equal(mug, muo)
        assert abs(mug.sum() - 1.0) < 1e-12
    mug, itg, resg = mg.power_iteration(100000, 1e-12)
    muo, ito, reso = mo.power_iteration(100000, 1e-12)
    assert resg <= 1e-12 and resg == reso and itg == ito
    assert np.array_equal(mug, muo)


def test_compress_index_vector():
    rng = np.random.default_rng(5)
    idx = np.sort(rng.integers(0, 1000, 20000)).astype(np.uint32)
    from oracle_api import compressIndexVector as ocv
    assert np.array_equal(gx.compressIndexVector(idx, 1000), ocv(idx, 1000))
    assert np.array_equal(gx.compressIndexVector(idx, 1000), np.searchsorted(idx, np.arange(1001)).astype(np.uint32))


def test_power_iteration_graph_replays_match_eager(monkeypatch):
    """Iterations of the batched form replayed from the captured HIP graph (many replays, host
    syncs between them) give the eager launches' and the oracle's measures bit for bit."""
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 11)
    m = g.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon())
    monkeypatch.setenv("B12X_PI_PERSISTENT", "0")
    mu_graph, it_graph, r_graph = m.power_iteration(700)
    monkeypatch.setenv("B12X_GRAPHS", "0")
    mu_eager, it_eager, r_eager = m.power_iteration(700)
    monkeypatch.delenv("B12X_GRAPHS")
    mu_again, it_again, _ = m.power_iteration(700)
    assert it_graph == it_eager == it_again == 700
    assert np.array_equal(mu_graph, mu_eager) and np.array_equal(mu_graph, mu_again) and r_graph == r_eager
    muo, ito, ro = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon()).power_iteration(700)
    assert ito == 700 and ro == r_graph and np.array_equal(mu_graph, muo)


def _pi_cases():
    """A subdivision matrix, an uploaded 300k-column stochastic matrix (a multi-group sweep per
    wave), a matrix with columns longer than the entries a lane keeps in registers, and a
    substochastic matrix whose norm decays by about 2^-10 per iteration (rescaled iterates)."""
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 13)
    cases = [("henon13", g.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon()),
              o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon()))]
    from oracle_api import OCoo
    rng = np.random.default_rng(11)
    n, per = 300_000, 4
    rows = np.repeat(np.arange(n, dtype=np.uint32), per)
    cols = rng.integers(0, n, n * per).astype(np.uint32)
    w = rng.random(n * per)
    vals = w / np.repeat(np.add.reduceat(w, np.arange(0, n * per, per)), per)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    keep = np.ones(len(rows), bool)
    keep[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    cases.append(("random300k", gx.CooMatrix.from_arrays(rows, cols, vals, n, n),
                  OCoo.from_arrays(rows, cols, vals, n, n)))
    # columns of 1 to 37 entries: the longer ones run past the entries a lane keeps in registers
    n = 4000
    lens = 1 + (np.arange(n) * 7919) % 37
    cols = np.repeat(np.arange(n, dtype=np.uint32), lens)
    rows = np.concatenate([rng.choice(n, k, replace=False) for k in lens]).astype(np.uint32)
    vals = rng.random(len(rows)) + 0.01
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    cases.append(("long-columns", gx.CooMatrix.from_arrays(rows, cols, vals, n, n),
                  OCoo.from_arrays(rows, cols, vals, n, n)))
    vals2 = vals / (1024.0 * np.bincount(rows, weights=vals, minlength=n)[rows])
    cases.append(("decaying", gx.CooMatrix.from_arrays(rows, cols, vals2, n, n),
                  OCoo.from_arrays(rows, cols, vals2, n, n)))
    # 20 uniformly spread rows per column, near the 24 entries a lane keeps
    n = 20_011
    cols = np.repeat(np.arange(n, dtype=np.uint32), 20)
    rows = np.concatenate([rng.choice(n, 20, replace=False) for _ in range(n)]).astype(np.uint32)
    vals = rng.random(len(rows)) + 0.01
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    vals = vals / np.bincount(rows, weights=vals, minlength=n)[rows]
    cases.append(("spread", gx.CooMatrix.from_arrays(rows, cols, vals, n, n),
                  OCoo.from_arrays(rows, cols, vals, n, n)))
    return cases


# the persistent form of each case
PI_FORMS = {"henon13": "straight", "random300k": "sweep", "long-columns": "sweep", "decaying": "sweep",
            "spread": "straight"}


def test_persistent_power_iteration_matches_batched_form_and_oracle(monkeypatch):
    """The single-launch (dataflow) power iteration against the batched step+reduce form and the
    oracle, bit for bit with equal iteration counts and residuals: iteration limits 1, 2, 3, 9,
    50 and a 1e-12 tolerance on every case of _pi_cases."""
    print("[pi] persistent vs batched vs oracle", flush=True)
    for name, mg, mo in _pi_cases():
        print(f"  case {name}", flush=True)
        for mi, tol in ((1, -1.0), (2, -1.0), (3, -1.0), (9, -1.0), (50, -1.0), (100000, 1e-12)):
            monkeypatch.setenv("B12X_PI_PERSISTENT", "2")  # persistent form or an error, no fallback
            a, ia, ra = mg.power_iteration(mi, tol)
            assert gx.power_iteration_kernel_ms() > 0.0
            assert gx.power_iteration_form() == PI_FORMS[name], name
            a2, ia2, ra2 = mg.power_iteration(mi, tol)
            monkeypatch.setenv("B12X_PI_PERSISTENT", "0")
            b, ib, rb = mg.power_iteration(mi, tol)
            assert gx.power_iteration_form() == "batched"
            monkeypatch.delenv("B12X_PI_PERSISTENT")
            c, ic, rc = mo.power_iteration(mi, tol)
            assert ia == ia2 == ib == ic and ra == ra2 == rb == rc, (name, mi, ia, ib, ic)
            assert np.array_equal(a, a2) and np.array_equal(a, b) and np.array_equal(a, c), (name, mi)
            if tol < 0:
                assert ia == mi
            else:
                assert ra <= tol


def test_persistent_power_iteration_gives_up_and_falls_back(monkeypatch):
    """With the persistent form's spin limit at 0 polls (set from the host), a wait gives up as
    soon as an entry or a total is not there at its first load: every wave leaves normally and the
    batched form recomputes the solve, equal bit for bit to the batched form alone."""
    print("[pi] spin limit 0: fallback", flush=True)
    g, o = subdivision_run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 16)
    m = g.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz())
    monkeypatch.setenv("B12X_PI_SPIN", "0")
    a, ia, ra = m.power_iteration(3000, 1e-12)
    fell_back = gx.power_iteration_kernel_ms() == 0.0
    monkeypatch.setenv("B12X_PI_PERSISTENT", "2")
    with pytest.raises(gx.B12xError, match="passed its limit"):
        m.power_iteration(3000, 1e-12)
    monkeypatch.delenv("B12X_PI_SPIN")
    monkeypatch.setenv("B12X_PI_PERSISTENT", "0")
    b, ib, rb = m.power_iteration(3000, 1e-12)
    monkeypatch.delenv("B12X_PI_PERSISTENT")
    c, ic, rc = m.power_iteration(3000, 1e-12)
    assert fell_back
    assert ia == ib == ic and ra == rb == rc
    assert np.array_equal(a, b) and np.array_equal(a, c)
    muo, ito, _ = o.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz()).power_iteration(3000, 1e-12)
    assert ito == ia and np.array_equal(a, muo)


def test_cpp_facade_doc_example(tmp_path):
    """B12's documented C++ example built against include/b12x/ImplicitBoxTree.h gives the
    oracle's leaf counts per step; the tree and the matrix go through their files and back."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "doc_example")
    assert os.path.exists(exe), "tests/cpp/doc_example not built (run __graft_entry__.build())"
    # 14 steps: depth 14 is a byte the reference's tree reader reads back (9-13 are whitespace)
    out = subprocess.run([exe, "14", str(tmp_path)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.splitlines()
    o = OracleTree([0.0, 0.0], [3.0, 3.0])
    o.setFlags(gx.SD, 0)
    counts = o.subdivision_steps(gx.InnerGrid(10), gx.Henon(1.4, 0.3, 1), 14)
    got = [int(line.split(":")[1]) for line in out[:14]]
    assert got == [int(c[3]) for c in counts]
    m = o.transitionMatrixForLeaves(gx.InnerGrid(10), gx.Henon(1.4, 0.3, 1))
    assert out[14] == f"nnz {m.nnz} rows {m.nRows} cols {m.nColumns}"
    mu, _, _ = m.power_iteration(50)
    assert float(out[15].split()[1]) == mu[0]
    assert out[16] == f"reloaded {int(counts[-1][3])} {m.nnz} 1"
    assert os.path.exists(tmp_path / "tree.json") and os.path.exists(tmp_path / "matrix.json")


@pytest.mark.parametrize("grid,use_invalid,row_major", [(gx.InnerGrid(10), False, False), (gx.InnerGrid(9), True, True),
                                                       (gx.UniformGrid(100, 3), False, True)])
def test_transition_matrix_more_than_64_points_per_box(grid, use_invalid, row_major):
    g, o = subdivision_run([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 8)
    mg = g.transitionMatrixForLeaves(grid, gx.Henon(), row_major=row_major, use_invalid=use_invalid)
    mo = o.transitionMatrixForLeaves(grid, gx.Henon(), row_major=row_major, use_invalid=use_invalid)
    assert (mg.nnz, mg.nRows, mg.nColumns) == (mo.nnz, mo.nRows, mo.nColumns)
    for a, b in zip(mg.download(), mo.download()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ppl,block,same_box", [(1, 64, 0), (2, 128, 0), (4, 256, 0), (4, 64, 0), (2, 256, 0),
                                                (2, 256, 1), (3, 128, 1), (3, 256, 1), (4, 128, 1)])
def test_image_kernel_variants_are_bit_exact(monkeypatch, ppl, block, same_box):
    """Every (points per lane, block size, same-box) variant the autotuner chooses between gives
    the oracle's covers, matrices and search results: Lorenz with InnerGrid (k = 27: the same-box
    form for 3 points a lane, points a block apart for 2 and 4), Hénon-Heiles with UniformGrid
    (k = 7: never the same-box form) on a random tree and with k = 64 on another."""
    print(f"[variant] ppl={ppl} block={block} same_box={same_box}", flush=True)
    monkeypatch.setenv("B12X_IMAGE_PPL", str(ppl))
    monkeypatch.setenv("B12X_IMAGE_BLOCK", str(block))
    monkeypatch.setenv("B12X_IMAGE_SAMEBOX", str(same_box))
    g, o = subdivision_run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 11)
    for a, b in zip(g.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz()).download(),
                    o.transitionMatrixForLeaves(gx.InnerGrid(3), gx.Lorenz()).download()):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(41)
    g2, o2 = both([0.0, 0.0, 0.0, 0.0], [0.6, 0.6, 0.5, 0.5])
    bits, depth, flags = random_tree(rng, 4, 900, max_depth=14)
    g2.upload(bits, depth, flags)
    o2.upload(bits, depth, flags)
    grid = gx.UniformGrid(7, 3)
    ig = g2.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
    io = o2.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
    assert np.array_equal(g2.searchByIterators(ig), o2.searchByIterators(io))
    grid = gx.UniformGrid(64, 0)
    ig = g2.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
    io = o2.makePointIteratorPairOverLeaves(grid, gx.HenonHeiles())
    assert np.array_equal(g2.searchByIterators(ig), o2.searchByIterators(io))
    assert g2.setFlagsByIterators(ig, 0x20) == o2.setFlagsByIterators(io, 0x20)
    assert_same_state(g2, o2)
