"""Insertion, deleteDepth and the continuation algorithm on the GPU against the oracle:
deleteDepth (ImplicitBoxTree.hpp:1217-1262), insertByPoints and insertByIterators
(ImplicitBoxTree.hpp:1613-1878) at depth 0, at leaf depth, at depths inside the tree, at the
deepest leaf's depth and beyond it, and the continuation driver (continuation.m). Trees and the
isToInsert arrays must be identical."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu


def pair(c, r, tree=None):
    g, o = gx.ImplicitBoxTree(c, r), OracleTree(c, r)
    if tree is not None:
        g.upload(*tree)
        o.upload(*tree)
    return g, o


def check(g, o, what):
    assert same_tree(g.download(), o.download()), what


def henon_fixed_point(a=1.4, b=0.3):
    x = (-(1 - b) + np.sqrt((1 - b) ** 2 + 4 * a)) / (2 * a)
    return np.array([[x, b * x]])


@pytest.mark.parametrize("dim,seed", [(1, 1), (2, 2), (3, 3), (4, 4)])
def test_delete_depth_matches_oracle(dim, seed):
    rng = np.random.default_rng(900 + seed)
    tree = random_tree(rng, dim, 3000, max_depth=18)
    c, r = rng.uniform(-2, 2, dim), rng.uniform(0.5, 3, dim)
    dmax = int(tree[1].max())
    for d in (1, 2, 5, 9, dmax - 1, dmax, dmax + 1, 254, 255, 0):
        print(f"[insert] deleteDepth dim={dim} seed={seed} d={d}", flush=True)
        g, o = pair(c, r, tree)
        assert g.deleteDepth(d) == o.deleteDepth(d), d
        check(g, o, d)
    # repeated deletion down to the root
    g, o = pair(c, r, tree)
    for d in range(dmax, 0, -1):
        assert g.deleteDepth(d) == o.deleteDepth(d)
        check(g, o, d)


@pytest.mark.parametrize("dim,seed", [(1, 11), (2, 12), (3, 13)])
def test_insert_by_points_matches_oracle(dim, seed):
    rng = np.random.default_rng(920 + seed)
    c, r = np.zeros(dim), np.ones(dim)
    tree = random_tree(rng, dim, 2000, max_depth=16)
    dmax = int(tree[1].max())
    for d in (0, 255, 1, 4, 9, 13, dmax, dmax + 3, 40):
        for fi, fh in ((0x02, 0x00), (0x20, 0x41)):
            print(f"[insert] byPoints dim={dim} seed={seed} d={d} flags={fi:#x},{fh:#x}", flush=True)
            g, o = pair(c, r, tree)
            pts = rng.uniform(-1.1, 1.1, (5000, dim))  # some outside the domain
            pts[:50] = pts[50:100]  # repeated points
            assert np.array_equal(g.insertByPoints(pts, d, fi, fh), o.insertByPoints(pts, d, fi, fh))
            check(g, o, (d, fi, fh))
            # a second insertion into the grown tree
            pts2 = rng.uniform(-1.0, 1.0, (3000, dim))
            assert np.array_equal(g.insertByPoints(pts2, d, fi, fh), o.insertByPoints(pts2, d, fi, fh))
            check(g, o, (d, fi, fh, 2))


def test_insert_by_points_into_the_surrounding_box_and_empty_input():
    """After deleteDepth(0) the surrounding box joins the run of the last lower-half and first
    upper-half boxes (the reference's merge, kept as it is); an empty point set changes nothing."""
    c, r = [0.0, 0.0], [1.0, 1.0]
    g, o = pair(c, r)
    for pts, d in (([[-0.5, 0.1], [0.3, 0.3], [0.31, 0.29]], 4), ([[0.2, -0.7]], 30),
                   ([[-0.5, -0.5], [-0.4, 0.6], [0.9, 0.9]], 255)):
        g.deleteDepth(0), o.deleteDepth(0)
        assert np.array_equal(g.insertByPoints(pts, d, 0x02), o.insertByPoints(pts, d, 0x02))
        check(g, o, d)
        before = g.download()
        assert g.insertByPoints(np.zeros((0, 2)), d, 0x02).size == 0
        assert same_tree(g.download(), before)


def test_insert_by_iterators_matches_oracle():
    c, r = [0.0, 0.0], [1.5, 0.5]
    grid, mp = gx.InnerGrid(4), gx.Henon()
    g, o = pair(c, r)
    g.setFlags(gx.SD, 0), o.setFlags(gx.SD, 0)
    g.subdivision_steps(grid, mp, 12)
    o.subdivision_steps(grid, mp, 12)
    check(g, o, "subdivision")
    rng = np.random.default_rng(931)
    st = rng.random(g.size()) < 0.3
    g.setFlagsByStencil(st, 0x10), o.setFlagsByStencil(st, 0x10)
    base = g.download()
    for it_depth, it_flags in ((255, 0), (255, 0x10), (8, 0), (6, 0x10), (0, 0)):
        for d in (255, 3, 8, 12, 15):
            print(f"[insert] byIterators it=({it_depth},{it_flags:#x}) d={d}", flush=True)
            g, o = pair(c, r, base)
            gi = g.makePointIteratorPairByFlags(it_depth, it_flags, grid, mp)
            oi = o.makePointIteratorPairByFlags(it_depth, it_flags, grid, mp)
            assert np.array_equal(g.insertByIterators(gi, d, 0x02, 0x01), o.insertByIterators(oi, d, 0x02, 0x01))
            check(g, o, (it_depth, it_flags, d))


def test_continuation_henon_matches_oracle():
    grid, mp = gx.InnerGrid(4), gx.Henon()
    p = henon_fixed_point()
    g, o = pair([0.0, 0.0], [3.0, 3.0])
    h = gx.ImplicitBoxTree([0.0, 0.0], [3.0, 3.0])
    for d in (8, 14, 20, 24):
        print(f"[insert] continuation henon d={d}", flush=True)
        cg = gx.continuation(g, grid, mp, [d], p)  # the loop on the device (b12x_continuation)
        co = gx.continuation(o, grid, mp, [d], p)
        assert cg == co, d
        check(g, o, d)
        steps = []  # the loop in Python over the GPU tree's methods
        assert gx.continuation(h, grid, mp, [d], p, on_step=lambda *x: steps.append(x)) == co
        assert same_tree(h.download(), o.download()) and len(steps) == len(co[0]) - 1


def test_continuation_lorenz_from_the_origin_matches_oracle():
    """The Lorenz system's unstable manifold of the origin, started from two points close to it
    on either side (one run per depth, trees compared after each)."""
    grid, mp = gx.InnerGrid(2), gx.Lorenz()
    c, r = [0.0, 0.0, 27.0], [30.0, 30.0, 40.0]
    p = np.array([[0.01, 0.01, 0.0], [-0.01, -0.01, 0.0]])
    g, o = pair(c, r)
    for d in (9, 15, 21):
        print(f"[insert] continuation lorenz d={d}", flush=True)
        cg = gx.continuation(g, grid, mp, [d], p)
        co = gx.continuation(o, grid, mp, [d], p)
        assert cg == co, d
        check(g, o, d)


def test_cpp_facade_continuation():
    """continuation.m written against include/b12x/ImplicitBoxTree.h (tests/cpp/continuation.cpp)
    prints the oracle's box counts per step."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "continuation")
    assert os.path.exists(exe), "tests/cpp/continuation not built (run __graft_entry__.build())"
    depths = [10, 16, 21]
    out = subprocess.run([exe] + [str(d) for d in depths], check=True, capture_output=True, text=True,
                         timeout=120).stdout.splitlines()
    o = OracleTree([0.0, 0.0], [3.0, 3.0])
    exp = gx.continuation(o, gx.InnerGrid(4), gx.Henon(), depths, henon_fixed_point())
    assert out == [f"{d}: " + " ".join(str(c) for c in cs) for d, cs in zip(depths, exp)]


@pytest.mark.parametrize("dim,sd,seed", [(2, [1, 0, 0], 41), (3, [2, 0, 1, 1], 42)])
def test_insertion_and_delete_depth_with_a_given_scheme(dim, sd, seed):
    """Trees with a given subdivision scheme: the boxes of inserted points, truncation by
    deleteDepth and the merge follow the scheme's bisection order."""
    rng = np.random.default_rng(960 + seed)
    c, r = rng.uniform(-1, 1, dim), rng.uniform(0.5, 2, dim)
    tree = random_tree(rng, dim, 1500, max_depth=15)
    g, o = gx.ImplicitBoxTree(c, r, sd), OracleTree(c, r, sd)
    g.upload(*tree)
    o.upload(*tree)
    dmax = int(tree[1].max())
    for d in (255, 5, dmax, dmax + 4):
        print(f"[insert] scheme {sd} byPoints d={d}", flush=True)
        pts = rng.uniform(c - 1.1 * r, c + 1.1 * r, (3000, dim))
        assert np.array_equal(g.insertByPoints(pts, d, 0x02, 0x40), o.insertByPoints(pts, d, 0x02, 0x40))
        check(g, o, (sd, d))
    for d in (9, 3):
        assert g.deleteDepth(d) == o.deleteDepth(d)
        check(g, o, (sd, "delete", d))


def test_continuation_and_chain_recurrent_set_with_a_given_scheme():
    grid, mp = gx.InnerGrid(4), gx.Henon()
    sd = [1, 0, 0]
    g, o = gx.ImplicitBoxTree([0.0, 0.0], [3.0, 3.0], sd), OracleTree([0.0, 0.0], [3.0, 3.0], sd)
    for d in (12, 20):
        print(f"[insert] continuation scheme {sd} d={d}", flush=True)
        assert gx.continuation(g, grid, mp, [d], henon_fixed_point()) == \
            gx.continuation(o, grid, mp, [d], henon_fixed_point())
        check(g, o, d)
    g, o = gx.ImplicitBoxTree([0.0, 0.0], [3.0, 3.0], sd), OracleTree([0.0, 0.0], [3.0, 3.0], sd)
    exp = gx.chainRecurrentSet(o, grid, mp, 20)
    g.setFlags(gx.SD, gx.LEAF)
    got = g.chain_recurrent_steps(grid, mp, 20)
    assert [tuple(int(v) for v in row) for row in got] == [tuple(e) for e in exp]
    check(g, o, "chain")


def test_fixed_boxes_matches_oracle():
    """fixedBoxes.m on the GPU tree and on the oracle: the same boxes after every number of steps
    up to the maximal depth."""
    for mp, grid, c, r in ((gx.Henon(), gx.InnerGrid(4), [0.0, 0.0], [3.0, 3.0]),
                           (gx.Henon(1.4, 0.3, 2), gx.InnerGrid(3), [0.0, 0.0], [3.0, 3.0]),
                           (gx.Lorenz(), gx.InnerGrid(2), [0.0, 0.0, 27.0], [30.0, 30.0, 40.0])):
        for steps in (12, 30, 64):
            print(f"[insert] fixedBoxes map={mp.kind} steps={steps}", flush=True)
            g, o = pair(c, r)
            bg = gx.fixedBoxes(g, grid, mp, steps)
            bo = gx.fixedBoxes(o, grid, mp, steps)
            assert np.array_equal(bg, bo)
            check(g, o, steps)


def test_subdivision_by_nim_matches_oracle():
    """subdivisionByNim.m with the power iteration as the eigenvector step: per step the same
    leaf counts (power iterations equal, or within one: the measure agrees within 1e-12), and the
    same trees."""
    for mp, grid, c, r, steps in ((gx.Henon(), gx.InnerGrid(4), [0.0, 0.0], [1.5, 0.5], 18),
                                  (gx.Lorenz(), gx.InnerGrid(3), [0.0, 0.0, 27.0], [30.0, 30.0, 30.0], 16)):
        print(f"[insert] subdivisionByNim map={mp.kind} steps={steps}", flush=True)
        g, o = pair(c, r)
        rg = gx.subdivisionByNim(g, grid, mp, steps)
        ro = gx.subdivisionByNim(o, grid, mp, steps)
        for a, b in zip(rg, ro):
            assert a[0] == b[0] and a[2] == b[2] and abs(a[1] - b[1]) <= 1
        check(g, o, steps)
