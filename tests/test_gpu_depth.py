"""GPU parity at depths between the root and the leaves and at the unsubdividable depth (254):
count, flag operations, searches, point iterators, flags by image and transition matrices, all
through the C ABI, bit for bit against the oracle on random trees with holes and mixed depths."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu

UNSUB = 254


def pair(rng, dim, n, max_depth):
    c, r = np.zeros(dim), np.ones(dim)
    g, o = gx.ImplicitBoxTree(c, r), OracleTree(c, r)
    bits, depth, flags = random_tree(rng, dim, n, max_depth=max_depth)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    return g, o


def depths_of(o):
    dmax = o.depth()
    return [d for d in range(1, dmax + 2)] + [UNSUB]


@pytest.mark.parametrize("dim,seed", [(2, 0), (3, 1), (2, 2)])
def test_count_and_flag_operations_at_depth(dim, seed):
    print(f"[depth] flags dim={dim} seed={seed}", flush=True)
    rng = np.random.default_rng(100 + seed)
    g, o = pair(rng, dim, 900, 13)
    for d in depths_of(o):
        assert g.count(d) == o.count(d), f"count({d})"
        n = o.count(d)
        st = rng.random(max(n, 1))[:n] < 0.5
        f = int(rng.choice([0x01, 0x08, 0x10, 0x20]))
        ops = [("setFlags", (f, d)), ("unsetFlags", (0x08, d)), ("changeFlags", (0x01, 0x40, d)),
               ("setFlagsByStencil", (st, f, d)), ("unsetFlagsByStencil", (st, 0x10, d)),
               ("changeFlagsByStencil", (st, 0x40, 0x02, d))]
        for name, args in ops:
            rg, ro = getattr(g, name)(*args), getattr(o, name)(*args)
            assert rg == ro, f"{name} at depth {d}: {rg} != {ro}"
            assert same_tree(g.download(), o.download()), f"{name} at depth {d}"


@pytest.mark.parametrize("dim,seed", [(2, 3), (3, 4)])
def test_search_by_points_at_depth(dim, seed):
    print(f"[depth] search dim={dim} seed={seed}", flush=True)
    rng = np.random.default_rng(200 + seed)
    g, o = pair(rng, dim, 1200, 12)
    pts = rng.uniform(-1.1, 1.1, (20000, dim))
    pts[:16] = 0.0  # box corners
    pts[16:32] = -1.0
    for d in depths_of(o):
        a, b = g.searchByPoints(pts, d), o.searchByPoints(pts, d)
        assert np.array_equal(a, b), f"searchByPoints at depth {d}"


@pytest.mark.parametrize("dim,seed", [(2, 5), (3, 6)])
def test_iterators_and_image_operations_at_depth(dim, seed):
    print(f"[depth] iterators dim={dim} seed={seed}", flush=True)
    rng = np.random.default_rng(300 + seed)
    g, o = pair(rng, dim, 700, 11)
    mp = gx.Henon(1.4, 0.3, 1) if dim == 2 else gx.Lorenz(steps=3)
    if dim == 3:  # the Lorenz flow over the unit cube: scale the domain to the attractor
        c, r = np.array([0.0, 0.0, 27.0]), np.array([30.0, 30.0, 30.0])
        bits, depth, flags = o.download()
        g, o = gx.ImplicitBoxTree(c, r), OracleTree(c, r)
        g.upload(bits, depth, flags)
        o.upload(bits, depth, flags)
    grid = gx.InnerGrid(2)
    for dIt in [3, 6, o.depth(), UNSUB]:
        for check in (0x00, 0x08, 0x09):
            ig = g.makePointIteratorPairByFlags(dIt, check, grid, mp)
            io = o.makePointIteratorPairByFlags(dIt, check, grid, mp)
            assert g.num_points(ig) == o.num_points(io), f"points at depth {dIt} check {check}"
            for d in (-1, 2, 5, UNSUB):
                assert np.array_equal(g.searchByIterators(ig, d), o.searchByIterators(io, d)), (dIt, check, d)
                assert np.array_equal(g.searchCompacted(ig, d), o.searchCompacted(io, d)), (dIt, check, d)
            for d in (4, UNSUB, -1):
                assert g.setFlagsByIterators(ig, 0x20, d) == o.setFlagsByIterators(io, 0x20, d)
                assert g.changeFlagsByIterators(ig, 0x20, 0x40, d) == o.changeFlagsByIterators(io, 0x20, 0x40, d)
                assert g.unsetFlagsByIterators(ig, 0x40, d) == o.unsetFlagsByIterators(io, 0x40, d)
                assert same_tree(g.download(), o.download()), (dIt, check, d)


@pytest.mark.parametrize("use_abs,row_major,use_invalid", [(False, False, False), (True, True, True),
                                                           (False, True, False)])
def test_transition_matrix_at_depth(use_abs, row_major, use_invalid):
    print(f"[depth] matrix abs={use_abs} rm={row_major} inv={use_invalid}", flush=True)
    rng = np.random.default_rng(400)
    c, r = np.array([0.0, 0.0]), np.array([1.5, 0.5])
    g, o = gx.ImplicitBoxTree(c, r), OracleTree(c, r)
    bits, depth, flags = random_tree(rng, 2, 1500, max_depth=14)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    grid, mp = gx.InnerGrid(3), gx.Henon(1.4, 0.3, 1)
    for d in (2, 5, 8, 11, o.depth() + 1, UNSUB):
        mg = g.transitionMatrixForDepth(d, grid, mp, use_abs, row_major, use_invalid)
        mo = o.transitionMatrixForDepth(d, grid, mp, use_abs, row_major, use_invalid)
        assert (mg.nnz, mg.nRows, mg.nColumns) == (mo.nnz, mo.nRows, mo.nColumns), f"depth {d}"
        for a, b in zip(mg.download(), mo.download()):
            assert np.array_equal(a, b), f"depth {d}"
    # iterator at one depth, columns at another
    for dIt, d in ((9, 4), (UNSUB, 6), (4, UNSUB)):
        ig = g.makePointIteratorPairOverDepth(dIt, grid, mp)
        io = o.makePointIteratorPairOverDepth(dIt, grid, mp)
        mg = g.transitionMatrix(ig, d, use_abs, row_major, use_invalid)
        mo = o.transitionMatrix(io, d, use_abs, row_major, use_invalid)
        assert (mg.nnz, mg.nRows, mg.nColumns) == (mo.nnz, mo.nRows, mo.nColumns), (dIt, d)
        for a, b in zip(mg.download(), mo.download()):
            assert np.array_equal(a, b), (dIt, d)


def test_unsubdividable_depth_count_matches_unsubdivide():
    """count(254) equals the leaf count unsubdivide(0) leaves behind, on the GPU as on the
    oracle, and both trees agree afterwards."""
    rng = np.random.default_rng(500)
    g, o = pair(rng, 2, 800, 10)
    n = g.count(UNSUB)
    assert n == o.count(UNSUB)
    assert g.unsubdivide(0) == o.unsubdivide(0) == n
    assert same_tree(g.download(), o.download())


@pytest.mark.parametrize("dim,sd,seed", [(1, None, 7), (2, None, 8), (3, None, 9), (4, None, 10), (2, [1, 0, 0, 1, 1], 11)])
def test_boxes_and_enumerate_match_oracle(dim, sd, seed):
    """boxes(depth) bit for bit (centres, radii, combined flags) at the root, every depth between
    root and leaves, one past the deepest leaf, the unsubdividable depth and the leaves; the
    general subdivision scheme included; enumerate() equal."""
    print(f"[depth] boxes dim={dim} sd={sd} seed={seed}", flush=True)
    rng = np.random.default_rng(600 + seed)
    c, r = rng.uniform(-1, 1, dim), rng.uniform(0.5, 2, dim)
    g, o = gx.ImplicitBoxTree(c, r, sd), OracleTree(c, r, sd)
    bits, depth, flags = random_tree(rng, dim, 700, max_depth=12)
    g.upload(bits, depth, flags)
    o.upload(bits, depth, flags)
    assert np.array_equal(g.enumerate(), o.enumerate())
    for d in list(range(0, o.depth() + 2)) + [UNSUB, -1]:
        a, b = g.boxes(d), o.boxes(d)
        assert a.shape == b.shape, d
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), d
