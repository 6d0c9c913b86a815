"""CPU tests (no GPU): pin the oracle against the reference's own code where it builds
(oracle/_ref), against library routines and golden vectors, and against closed forms,
invariants and brute force; check that the HIP library loads and exports its C ABI."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree
from oracle_api import OCoo, OracleTree, compressIndexVector, ref_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
P = ctypes.c_void_p
D = ctypes.c_double


def oscheme(dim, grid, mp):
    return OracleTree(np.zeros(dim), np.ones(dim)).scheme(grid, mp)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


# ------------------------------------------------------------------ integrators vs reference code
@pytest.fixture(scope="module")
def ref():
    r = ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return r


def test_lorenz_rk4auto_matches_reference_code(ref):
    rng = np.random.default_rng(0)
    x = rng.uniform(-25, 25, (3000, 3)) + [0, 0, 27]
    for sig, rho, beta, h, st in [(10, 28, 8 / 3, 0.01, 20), (10, 24.3, 8 / 3, 0.01, 20), (9.5, 31, 2.5, 0.005, 7)]:
        y = oscheme(3, gx.InnerGrid(3), gx.Lorenz(sig, rho, beta, h, st)).map_points(x)
        z = np.empty_like(x)
        ref.ref_lorenz(D(sig), D(rho), D(beta), D(h), ctypes.c_uint32(st), x.ctypes.data_as(P), z.ctypes.data_as(P),
                       ctypes.c_uint64(len(x)))
        assert bits_equal(y, z)


def test_henon_iterating_map_matches_reference_code(ref):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1.5, 1.5, (3000, 2))
    for a, b, st in [(1.4, 0.3, 1), (1.4, 0.3, 2), (1.2, 0.3, 5), (1.0, 0.3, 0)]:
        y = oscheme(2, gx.InnerGrid(4), gx.Henon(a, b, st)).map_points(x)
        z = np.empty_like(x)
        ref.ref_henon(D(a), D(b), ctypes.c_uint32(st), x.ctypes.data_as(P), z.ctypes.data_as(P),
                      ctypes.c_uint64(len(x)))
        assert bits_equal(y, z)


def test_duffing_rk4_matches_reference_code(ref):
    rng = np.random.default_rng(2)
    x = rng.uniform(-2, 2, (2000, 2))
    for t1, h in [(2 * np.pi, 2 * np.pi / 100), (1.0, 0.3), (2.0, 0.7)]:
        mp = gx.Duffing(0.25, -1.0, 1.0, 0.3, 1.0, 0.0, t1, h)
        y = oscheme(2, gx.InnerGrid(4), mp).map_points(x)
        z = np.empty_like(x)
        ref.ref_duffing(D(0.25), D(-1.0), D(1.0), D(0.3), D(1.0), D(0.0), D(t1), D(h), x.ctypes.data_as(P),
                        z.ctypes.data_as(P), ctypes.c_uint64(len(x)))
        assert bits_equal(y, z)


def test_duffing_step_count_follows_float_time_loop():
    """The non-autonomous RK4 loop counts steps with a float time variable (MapTools.hpp:54):
    T = 1.0 with h = 0.1 accumulates to 0.9999999999999999 < 1 and takes an 11th tiny step."""
    t, n = 0.0, 0
    while t < 1.0:
        t += 0.1
        n += 1
    assert n == 11
    x = np.array([[0.3, -0.2]])
    full = oscheme(2, gx.InnerGrid(1), gx.Duffing(0.25, -1, 1, 0.3, 1, 0.0, 1.0, 0.1)).map_points(x)
    # the 11th step has h ~ 1e-16: result equals 10 steps up to that tiny step
    ten = oscheme(2, gx.InnerGrid(1), gx.Duffing(0.25, -1, 1, 0.3, 1, 0.0, 0.9999999999999999, 0.1)).map_points(x)
    assert np.allclose(full, ten, rtol=0, atol=1e-14)


# ------------------------------------------------------------------ integrator invariants
def test_rk4_order_four_convergence():
    x = np.array([[1.0, 2.0, 20.0]])
    T = 0.2
    sols = [oscheme(3, gx.InnerGrid(1), gx.Lorenz(h=T / n, steps=n)).map_points(x) for n in (10, 20, 40, 80)]
    e = [np.abs(sols[i] - sols[i + 1]).max() for i in range(3)]
    ratios = [e[i] / e[i + 1] for i in range(2)]
    assert all(12 < r < 20 for r in ratios), ratios  # 2^4 = 16


def hh_energy(z):
    q1, q2, p1, p2 = z.T
    return 0.5 * (p1 ** 2 + p2 ** 2) + 0.5 * (q1 ** 2 + q2 ** 2) + q1 ** 2 * q2 - q2 ** 3 / 3


@pytest.mark.parametrize("order,expected", [(2, 4.0), (4, 16.0)])
def test_symplectic_energy_error_order(order, expected):
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.3, 0.3, (200, 4))
    errs = []
    for h, st in [(0.04, 25), (0.02, 50), (0.01, 100)]:
        y = oscheme(4, gx.InnerGrid(1), gx.HenonHeiles(order, h, st)).map_points(x)
        errs.append(np.abs(hh_energy(y) - hh_energy(x)).max())
    r = errs[0] / errs[1], errs[1] / errs[2]
    assert all(0.6 * expected < v < 1.6 * expected for v in r), r


def test_stormer_verlet_time_reversible():
    rng = np.random.default_rng(4)
    x = rng.uniform(-0.3, 0.3, (100, 4))
    s = oscheme(4, gx.InnerGrid(1), gx.HenonHeiles(2, 0.02, 50))
    y = s.map_points(x)
    y[:, 2:] *= -1
    z = s.map_points(y)
    z[:, 2:] *= -1
    assert np.abs(z - x).max() < 1e-13


def test_henon_closed_form():
    x = np.array([[0.5, 0.25], [-1.0, 0.1]])
    y = oscheme(2, gx.InnerGrid(1), gx.Henon(1.4, 0.3, 1)).map_points(x)
    assert bits_equal(y, np.array([[1.0 - 1.4 * 0.5 * 0.5 + 0.25, 0.3 * 0.5], [1.0 - 1.4 * -1.0 * -1.0 + 0.1, -0.3]]))


# ------------------------------------------------------------------ UniformGrid vs libstdc++
def mt19937_python(seed):
    """Independent MT19937 (Matsumoto & Nishimura 1998) for the KAT."""
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = 624
    while True:
        if idx >= 624:
            for k in range(624):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        yield y & 0xFFFFFFFF


def test_mt19937_known_answer():
    g = mt19937_python(5489)
    v = None
    for _ in range(10000):
        v = next(g)
    golden = json.load(open(os.path.join(GOLDEN, "uniform_grid.json")))
    assert v == 4123659995 == golden["mt19937_default_10000th"]


def test_uniform_grid_matches_libstdcxx_golden():
    golden = json.load(open(os.path.join(GOLDEN, "uniform_grid.json")))
    for row in golden["rows"]:
        s = oscheme(4, gx.UniformGrid(64, row["seed"]), gx.Identity(4))
        got = s.grid_points(row["i"], 1)[0]
        want = np.array([float.fromhex(h) for h in row["draws_hex"]])
        assert bits_equal(got, want), row


# ------------------------------------------------------------------ keys and descent
def full_tree(dim, d, c, r):
    o = OracleTree(c, r)
    o.setFlags(gx.SD, 0)
    for _ in range(d):
        o.subdivide(gx.SD)
    return o


@pytest.mark.parametrize("dim,d", [(1, 7), (2, 8), (3, 9), (4, 8)])
def test_descent_brute_force_geometry(dim, d):
    """On a complete tree of depth d, the leaf found for a point is the box that contains it,
    with the box bounds recomputed independently from the leaf's bit pattern."""
    rng = np.random.default_rng(d)
    c = rng.uniform(-1, 1, dim)
    r = rng.uniform(0.5, 3, dim)
    o = full_tree(dim, d, c, r)
    bits, depth, _ = o.download()
    assert o.size() == 2 ** d and np.all(depth == d)
    pts = rng.uniform(c - r, c + r, (3000, dim))
    idx = o.searchByPoints(pts)
    assert np.all(idx < 2 ** d)
    # independent geometry: level i bisects coordinate i % dim; bit (d-1-i) is the side
    for p, i in zip(pts, idx):
        b = int(bits[i])
        lo, hi = c - r, c + r
        lo, hi = lo.copy(), hi.copy()
        for lev in range(d):
            j = lev % dim
            mid = 0.5 * (lo[j] + hi[j])
            if (b >> (d - 1 - lev)) & 1:
                lo[j] = mid
            else:
                hi[j] = mid
        assert np.all(p >= lo - 1e-12) and np.all(p <= hi + 1e-12)
    # outside the domain -> invalid
    out = o.searchByPoints(np.array([c + 2 * r, c - 2 * r]))
    assert np.all(out == gx.INVALID)


def test_descent_random_trees_brute_force():
    rng = np.random.default_rng(9)
    for trial in range(5):
        o = OracleTree([0.0, 0.0], [1.0, 1.0])
        bits, depth, flags = random_tree(rng, 2, 400, max_depth=10)
        o.upload(bits, depth, flags)
        b, d, _ = o.download()
        pts = rng.uniform(-1, 1, (2000, 2))
        idx = o.searchByPoints(pts)
        full = full_tree(2, 10, [0.0, 0.0], [1.0, 1.0])
        fb, fd, _ = full.download()
        cell = fb[full.searchByPoints(pts)]  # depth-10 cell of each point
        for ci, i in zip(cell, idx):
            owners = [k for k in range(len(b)) if (int(ci) >> (10 - int(d[k]))) == int(b[k])]
            assert len(owners) <= 1
            assert (i == gx.INVALID and not owners) or (owners and i == owners[0])


def test_sorted_by_comparable_key():
    rng = np.random.default_rng(5)
    o = OracleTree([0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    o.upload(*random_tree(rng, 3, 500, max_depth=12))
    b, d, _ = o.download()
    key = [(int(x) << (64 - int(dd))) | ((1 << (63 - int(dd))) - 1) if dd < 64 else int(x) for x, dd in zip(b, d)]
    assert key == sorted(key) and len(set(key)) == len(key)


# ------------------------------------------------------------------ tree invariants
def test_subdivide_unsubdivide_roundtrip():
    rng = np.random.default_rng(6)
    for dim in (1, 2, 3, 4):
        o = OracleTree(np.zeros(dim), np.ones(dim))
        bits, depth, _ = random_tree(rng, dim, 200, max_depth=9)
        o.upload(bits, depth, np.full(len(bits), gx.SD, np.uint8))
        before = o.download()
        n = o.size()
        assert o.subdivide(gx.SD) == n
        assert o.size() == 2 * n
        assert o.unsubdivide(gx.SD) == n
        assert same_tree(o.download(), before)


def _unsubdivide_brute_force(bits, depth, flags, f):
    """unsubdivide(f) restated on bit patterns alone (Functors.hpp:337-377, ImplicitBoxTree.hpp:
    2023-2058,3413-3461): a leaf of depth > 0 with all of f set merges into its parent iff its
    sibling (the parent's other half) is a leaf with all of f set, or no leaf lies inside the
    sibling's region at all; equal parents become one box with the AND of their flags. Returns
    (#parents formed, the resulting leaf set as {(bits, depth): flags})."""
    leaves = {(int(b), int(d)): int(fl) for b, d, fl in zip(bits, depth, flags)}

    def inside(region, leaf):  # leaf lies in the region (region is an ancestor of it or it)
        (rb, rd), (lb, ld) = region, leaf
        return ld >= rd and (lb >> (ld - rd)) == rb

    out, parents = {}, {}
    for (b, d), fl in leaves.items():
        merge = False
        if d > 0 and (fl & f) == f:
            sib = (b ^ 1, d)
            if sib in leaves:
                merge = (leaves[sib] & f) == f
            else:
                merge = not any(inside(sib, l) for l in leaves)
        if merge:
            p = (b >> 1, d - 1)
            parents[p] = parents.get(p, 0xFF) & fl
        else:
            out[(b, d)] = fl
    assert not set(parents) & set(out)  # a formed parent never coincides with a kept leaf
    out.update(parents)
    return len(parents), out


def test_unsubdivide_merge_rule_brute_force():
    """unsubdivide on random trees with holes, unflagged siblings, deeper siblings and mixed
    flags against the bit-pattern restatement above, for several flag masks."""
    rng = np.random.default_rng(61)
    cases = 0
    for trial in range(40):
        dim = 1 + trial % 4
        bits, depth, flags = random_tree(rng, dim, int(rng.integers(2, 160)), max_depth=int(rng.integers(1, 9)),
                                         flag_prob=0.7, flags_pool=(0x08, 0x01, 0x09, 0x18, 0x00))
        for f in (gx.SD, gx.HIT, gx.SD | gx.HIT, 0):
            o = OracleTree(np.zeros(dim), np.ones(dim))
            o.upload(bits, depth, flags)
            res, expect = _unsubdivide_brute_force(*o.download(), f)
            assert o.unsubdivide(f) == res
            b, d, fl = o.download()
            assert {(int(x), int(y)): int(z) for x, y, z in zip(b, d, fl)} == expect
            cases += res > 0
    assert cases > 40  # most cases merge something


def _leaf_index_of_points(bits, depth, pts, c, r):
    """Index of the leaf holding each point by its depth-64 key (coordinateFunctions.hpp:61-93:
    u = ((p-c)/r + 1)/2, b_j = floor(2^nb_j u), bit i of b_j at position i*dim + (63-j)%dim),
    or INVALID outside the domain or in a hole; brute force over the leaves."""
    dim = len(c)
    nb = [64 // dim + (j < 64 % dim) for j in range(dim)]
    out = np.full(len(pts), gx.INVALID, np.uint32)
    for q, p in enumerate(pts):
        if not np.all((c - r <= p) & (p < c + r)):
            continue
        key = 0
        for j in range(dim):
            u = ((p[j] - c[j]) / r[j] + 1.0) / 2.0
            bj = int(np.floor(np.ldexp(u, nb[j])))
            for i in range(nb[j]):
                pos = i * dim + (63 - j) % dim
                if pos < 64 and (bj >> i) & 1:
                    key |= 1 << pos
        for k, (b, d) in enumerate(zip(bits, depth)):
            if (key >> (64 - int(d))) == int(b) if int(d) > 0 else True:
                out[q] = k
                break
    return out


def _grid_points_of_leaf(b, d, c, r, n_per_dim):
    """InnerGrid(n) points of leaf (b, d), default scheme (Grids.hpp:99-134, coordinateFunctions.
    hpp:128-159): level l bisects coordinate l % dim; centre ((2^-nb idx)*2 - 1)*r + c + r_box."""
    dim = len(c)
    idx = [0] * dim
    nbj = [0] * dim
    for lev in range(d):
        j = lev % dim
        idx[j] = (idx[j] << 1) | ((b >> (d - 1 - lev)) & 1)
        nbj[j] += 1
    rb = np.array([np.ldexp(r[j], -nbj[j]) for j in range(dim)])
    cb = np.array([((np.ldexp(float(idx[j]), -nbj[j]) * 2.0 - 1.0) * r[j] + c[j]) + rb[j] for j in range(dim)])
    k = n_per_dim ** dim
    pts = np.empty((k, dim))
    for i in range(k):
        for j in range(dim):
            g = ((float((i // n_per_dim ** j) % n_per_dim) * 2.0 + 1.0) - n_per_dim) / n_per_dim
            pts[i, j] = cb[j] + g * rb[j]
    return pts


@pytest.mark.parametrize("row_major", [False, True])
@pytest.mark.parametrize("use_abs", [False, True])
def test_transition_matrix_invalid_column_label_hand_built_tree(row_major, use_abs):
    """transitionMatrix with useInvalidNrBoxes on a hand-built Hénon tree with holes, recomputed
    from the grid formula, the map's closed form and brute-force leaf geometry: with invalid
    targets the matrix has n + 1 columns and they are labelled n + 1, nColumns_ after its
    increment (ImplicitBoxTree.hpp:3081-3087 row-major, :3127-3132 column-major); without the
    flag they are dropped."""
    c, r = np.array([0.0, 0.0]), np.array([1.5, 0.5])
    # the depth-4 boxes of the domain, with holes
    keep = [0, 1, 2, 3, 5, 6, 9, 10, 11, 12, 13, 15]
    bits = np.array(keep, np.uint64)
    depth = np.full(len(keep), 4, np.uint8)
    o = OracleTree(c, r)
    o.upload(bits, depth, np.zeros(len(keep), np.uint8))
    b, d, _ = o.download()
    n, kpd = len(b), 2
    k = kpd ** 2
    rows, cols = [], []
    for i in range(n):
        x = _grid_points_of_leaf(int(b[i]), int(d[i]), c, r, kpd)
        y = np.stack([(1.0 - (1.4 * x[:, 0]) * x[:, 0]) + x[:, 1], 0.3 * x[:, 0]], axis=1)
        tgt = _leaf_index_of_points(b, d, y, c, r)
        rows += [i] * k
        cols += list(tgt)
    rows, cols = np.array(rows, np.int64), np.array(cols, np.int64)
    assert (cols == gx.INVALID).sum() > 0 and (cols != gx.INVALID).sum() > 0
    for use_invalid in (False, True):
        m = o.transitionMatrixForLeaves(gx.InnerGrid(kpd), gx.Henon(), use_abs=use_abs, row_major=row_major,
                                        use_invalid=use_invalid)
        pairs = {}
        for a, z in zip(rows, cols):
            if z == gx.INVALID and not use_invalid:
                continue
            pairs[(a, z)] = pairs.get((a, z), 0) + 1
        # sort with the invalid label (0xFFFFFFFF) last, as before the relabelling, then relabel
        order = sorted(pairs, key=(lambda e: (e[0], e[1])) if row_major else (lambda e: (e[1], e[0])))
        er = np.array([e[0] for e in order], np.uint32)
        ec = np.array([n + 1 if e[1] == gx.INVALID else e[1] for e in order], np.uint32)
        ev = np.array([pairs[e] if use_abs else pairs[e] / float(k) for e in order], np.float64)
        gr, gc, gv = m.download()
        assert m.nRows == n and m.nColumns == (n + 1 if use_invalid else n)
        assert np.array_equal(gr, er) and np.array_equal(gc, ec) and bits_equal(gv, ev)
        if use_invalid:
            assert np.sum(gc == n + 1) > 0 and np.all(gc != n)


def test_remove_keeps_flagged_and_reinserts_root():
    o = OracleTree([0.0, 0.0], [1.0, 1.0])
    o.setFlags(gx.SD, 0)
    for _ in range(4):
        o.subdivide(gx.SD)
    st = np.arange(16) % 3 == 0
    assert o.setFlagsByStencil(st, gx.HIT) == st.sum()
    assert o.remove(gx.HIT) == 16 - st.sum()
    assert o.size() == st.sum()
    assert o.unsetFlags(gx.HIT) == st.sum()
    assert o.remove(gx.HIT) == st.sum()
    b, d, f = o.download()
    assert list(b) == [0] and list(d) == [0] and list(f) == [0]


def test_flag_return_values():
    o = OracleTree([0.0], [1.0])
    o.upload(np.array([0, 1, 2, 3], np.uint64), np.full(4, 2, np.uint8), np.array([0, 1, 3, 2], np.uint8))
    assert o.setFlags(1) == 2          # leaves lacking HIT
    assert o.unsetFlags(2) == 2        # leaves with any of flags
    assert o.changeFlags(1, 4) == 4    # leaves with all of flagsToUnset
    assert o.count(-1) == 4 and o.count(2) == 4 and o.count(0) == 1 and o.count(1) == 2 and o.count(3) == 0
    assert o.depth() == 2


def test_identity_map_hits_every_leaf_and_gives_identity_matrix():
    rng = np.random.default_rng(7)
    o = OracleTree([0.0, 0.0, 0.0], [1.0, 2.0, 3.0])
    o.upload(*random_tree(rng, 3, 300, max_depth=9))
    it = o.makePointIteratorPairOverLeaves(gx.InnerGrid(2), gx.Identity(3))
    assert np.array_equal(o.searchCompacted(it), np.arange(o.size()))
    m = o.transitionMatrixForLeaves(gx.InnerGrid(2), gx.Identity(3))
    r, c, v = m.download()
    assert np.array_equal(r, np.arange(o.size())) and np.array_equal(c, r) and np.all(v == 1.0)


def test_transition_matrix_counts_and_orderings():
    """Row sums equal the fraction of a box's points landing inside the cover; the column-major
    and row-major matrices hold the same entries; absolute values are integer counts."""
    o = OracleTree([0.0, 0.0], [1.5, 0.5])
    o.setFlags(gx.SD, 0)
    o.subdivision_steps(gx.InnerGrid(4), gx.Henon(), 8)
    it = o.makePointIteratorPairOverLeaves(gx.InnerGrid(4), gx.Henon())
    idx = o.searchByIterators(it).reshape(-1, 16)
    inside = (idx != gx.INVALID).sum(axis=1)
    mc = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon())
    mr = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), row_major=True)
    ma = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), use_abs=True, row_major=True)
    rc, cc, vc = mc.download()
    rr, cr, vr = mr.download()
    ra, ca, va = ma.download()
    assert np.all(np.diff(cc.astype(np.int64) * 2 ** 32 + rc) > 0)
    assert np.all(np.diff(rr.astype(np.int64) * 2 ** 32 + cr) > 0)
    sums = np.bincount(rr, weights=vr, minlength=o.size())
    assert np.allclose(sums, inside / 16.0)
    assert np.array_equal(va * 16 / 16, va) and np.all(va == np.round(va))
    assert np.allclose(np.bincount(ra, weights=va, minlength=o.size()), inside)
    o1 = np.lexsort((rc, cc))
    assert set(zip(rr, cr, vr)) == set(zip(rc, cc, vc)) and len(o1) == len(rr)


def test_search_compacted_chunk_invariance():
    o = OracleTree([0.0, 0.0, 27.0], [30.0, 30.0, 30.0])
    o.setFlags(gx.SD, 0)
    o.subdivision_steps(gx.InnerGrid(3), gx.Lorenz(), 9)
    it = o.makePointIteratorPairOverLeaves(gx.InnerGrid(3), gx.Lorenz(rho=25.0))
    full = o.searchCompacted(it)
    for chunk in (1, 7, 1000):
        o.set_options(False, chunk)
        assert np.array_equal(o.searchCompacted(it), full)
    o.set_options(True, 0)
    assert np.array_equal(o.searchCompacted(it), full)


def test_compress_index_vector_closed_form():
    rng = np.random.default_rng(8)
    idx = np.sort(rng.integers(0, 300, 5000)).astype(np.uint32)
    assert np.array_equal(compressIndexVector(idx, 300), np.searchsorted(idx, np.arange(301)).astype(np.uint32))
    assert np.array_equal(compressIndexVector(np.array([], np.uint32), 3), np.zeros(4, np.uint32))


# ------------------------------------------------------------------ invariant measure
def test_power_iteration_closed_forms():
    # doubly stochastic -> uniform
    n = 5
    P = np.full((n, n), 1.0 / n)
    r, c = np.nonzero(P)
    m = OCoo.from_arrays(r, c, P[r, c], n, n)
    mu, it, res = m.power_iteration(10)
    assert np.allclose(mu, 1.0 / n, rtol=1e-15)
    # 2-state chain: stationary distribution (b, a)/(a+b)
    a, b = 0.3, 0.1
    P = np.array([[1 - a, a], [b, 1 - b]])
    r, c = np.nonzero(P)
    mu, it, res = OCoo.from_arrays(r, c, P[r, c], 2, 2).power_iteration(100000, 1e-15)
    assert np.allclose(mu, [b / (a + b), a / (a + b)], rtol=1e-13)
    # a cyclic permutation keeps the uniform start fixed
    perm = np.roll(np.arange(7), 1)
    mu, _, res = OCoo.from_arrays(np.arange(7), perm, np.ones(7), 7, 7).power_iteration(3)
    assert np.allclose(mu, 1 / 7, rtol=1e-15) and res < 1e-15


def test_power_iteration_matches_numpy_eig():
    rng = np.random.default_rng(10)
    n = 40
    A = rng.random((n, n)) * (rng.random((n, n)) < 0.3) + np.eye(n) * 0.1
    P = A / A.sum(axis=1, keepdims=True) * 0.97  # substochastic, as when mass leaves the cover
    r, c = np.nonzero(P)
    mu, it, res = OCoo.from_arrays(r, c, P[r, c], n, n).power_iteration(100000, 1e-14)
    w, V = np.linalg.eig(P.T)
    k = np.argmax(np.abs(w))
    v = np.abs(V[:, k].real)
    v /= v.sum()
    assert np.allclose(mu, v, rtol=1e-9, atol=1e-13)


def test_power_iteration_exact_on_integer_iterates():
    """With integer counts (useAbsoluteValues) and an integer start vector every iterate z_t =
    z_0 P^t is an exact integer below 2^53, whatever the order of the sums, so mu_t = z_t / |z_t|_1
    is one correctly rounded division per entry: recomputed exactly in Python integers."""
    rng = np.random.default_rng(12)
    n = 300
    nnz = 1200
    rows = rng.integers(0, n, nnz)
    cols = rng.integers(0, n, nnz)
    key = np.unique(rows * n + cols)
    rows, cols = (key // n).astype(np.uint32), (key % n).astype(np.uint32)
    vals = rng.integers(1, 4, len(rows)).astype(np.float64)
    z0 = rng.integers(1, 5, n).astype(np.float64)
    for iters in (1, 2, 5, 7):
        mu, it, _ = OCoo.from_arrays(rows, cols, vals, n, n).power_iteration(iters, -1.0, mu=z0)
        z = [int(v) for v in z0]
        for _ in range(iters):
            nz = [0] * n
            for a, b, v in zip(rows.tolist(), cols.tolist(), vals.tolist()):
                nz[b] += z[a] * int(v)
            z = nz
        assert max(z) < 2 ** 53
        total = sum(abs(v) for v in z)
        assert it == iters and bits_equal(mu, np.array([float(v) / float(total) for v in z]))


def test_power_iteration_rescales_decaying_iterates():
    """A matrix whose rows sum to 2^-10: unscaled iterates would fall below the FP64 range after
    about 110 iterations. The oracle's exact power-of-two rescaling keeps them in range, and its
    measure after 400 iterations equals numpy's normalised iteration mu <- mu P / |mu P|_1."""
    rng = np.random.default_rng(13)
    n = 60
    A = rng.random((n, n)) * (rng.random((n, n)) < 0.4) + np.eye(n)
    P = A / A.sum(axis=1, keepdims=True) * 2.0 ** -10
    r, c = np.nonzero(P)
    mu, it, res = OCoo.from_arrays(r, c, P[r, c], n, n).power_iteration(400)
    v = np.full(n, 1.0 / n)
    for _ in range(400):
        v = v @ P
        v /= v.sum()
    assert it == 400 and np.all(np.isfinite(mu)) and res < 1e-12
    assert np.allclose(mu, v, rtol=1e-12, atol=0.0)


# ------------------------------------------------------------------ depths between root and leaves
def depth_groups(bits, depth, d):
    """Boxes at depth d, brute force: the ancestors at d of the leaves of depth >= d, in leaf
    (key) order, and for every leaf its group index (-1 for leaves shallower than d)."""
    groups, gid = [], []
    for b, dd in zip(bits.tolist(), depth.tolist()):
        if dd < d:
            gid.append(-1)
            continue
        a = b >> (dd - d)
        if not groups or groups[-1] != a:
            groups.append(a)
        gid.append(len(groups) - 1)
    return groups, np.array(gid)


def general_tree(rng, dim=2, n=700, max_depth=14):
    o = OracleTree([0.0] * dim, [1.0] * dim)
    o.upload(*random_tree(rng, dim, n, max_depth=max_depth))
    return o


def test_count_and_flags_at_depths_between_root_and_leaves():
    """count(d), and set/unset/changeFlags(ByStencil) at depth d against a brute-force grouping
    of the leaves by their depth-d ancestor (flags go to every leaf of a selected box; the
    return value counts the selected boxes whose leaves fail the predicate jointly)."""
    rng = np.random.default_rng(51)
    for trial in range(4):
        o = general_tree(rng)
        bits, depth, flags = o.download()
        for d in range(1, int(depth.max()) + 1):
            if np.all(depth == d):
                continue
            groups, gid = depth_groups(bits, depth, d)
            assert o.count(d) == len(groups)
            st = rng.random(len(groups)) < 0.5
            f = int(rng.choice([0x01, 0x08, 0x10]))
            before = o.download()[2].copy()
            ret = o.setFlagsByStencil(st.astype(np.uint8), f, d)
            after = o.download()[2]
            sel = np.array([g >= 0 and st[g] for g in gid])
            assert np.array_equal(after, np.where(sel, before | f, before))
            exp = sum(1 for g in range(len(groups)) if st[g] and not np.all((before[gid == g] & f) == f))
            assert ret == exp
            ret = o.unsetFlagsByStencil(st.astype(np.uint8), f, d)
            exp = sum(1 for g in range(len(groups)) if st[g] and np.any(after[gid == g] & f))
            assert ret == exp
            assert np.array_equal(o.download()[2], np.where(sel, after & np.uint8(0xFF ^ f), after))
            o.upload(bits, depth, flags)


def test_search_at_depth_maps_points_to_their_depth_box():
    """A point is found in the box at depth d containing it iff that box holds a leaf of depth
    >= d (a hole inside the box still finds it); otherwise invalid."""
    rng = np.random.default_rng(52)
    o = general_tree(rng, 2, 900, 12)
    bits, depth, _ = o.download()
    pts = rng.uniform(-1.2, 1.2, (3000, 2))
    for d in (1, 3, 5, 8):
        groups, _ = depth_groups(bits, depth, d)
        pos = {a: i for i, a in enumerate(groups)}
        got = o.searchByPoints(pts, d)
        for p, g in zip(pts, got):
            if np.any(p < -1.0) or np.any(p >= 1.0):
                assert g == gx.INVALID
                continue
            a = 0  # depth-d ancestor of the point: bisect x then y, alternately
            lo, hi = np.array([-1.0, -1.0]), np.array([1.0, 1.0])
            for lvl in range(d):
                j = lvl % 2
                mid = 0.5 * (lo[j] + hi[j])
                bit = 1 if p[j] >= mid else 0
                a = (a << 1) | bit
                if bit:
                    lo[j] = mid
                else:
                    hi[j] = mid
            assert g == pos.get(a, gx.INVALID)


def test_identity_matrix_at_depth_and_unsubdividable_count():
    rng = np.random.default_rng(53)
    o = general_tree(rng, 2, 500, 10)
    bits, depth, flags = o.download()
    for d in (2, 4, 6):
        m = o.transitionMatrixForDepth(d, gx.InnerGrid(3), gx.Identity(2))
        r, c, v = m.download()
        n = o.count(d)
        assert m.nRows == m.nColumns == n
        assert np.array_equal(r, np.arange(n)) and np.array_equal(c, np.arange(n)) and np.all(v == 1.0)
    n_unsub = o.count(254)
    assert n_unsub == o.unsubdivide(0)


def bisect_box(b, d, c, r):
    """Bounds of the box with pattern b at depth d (default scheme: coordinate lvl % dim),
    by bisecting the domain level by level (independent of the oracle)."""
    dim = len(c)
    lo, hi = np.array(c, float) - np.array(r, float), np.array(c, float) + np.array(r, float)
    for lvl in range(d):
        j = lvl % dim
        mid = 0.5 * (lo[j] + hi[j])
        if (b >> (d - 1 - lvl)) & 1:
            lo[j] = mid
        else:
            hi[j] = mid
    return lo, hi


def test_boxes_and_enumerate_against_bisection():
    """boxes(depth): centre and radius from bisecting the domain along the pattern, flags ANDed
    over the leaves of each box (the root: over all leaves), boxes in leaf order, nothing beyond
    the deepest leaf; enumerate() = the number of distinct depth-d ancestors per depth."""
    rng = np.random.default_rng(61)
    for dim in (2, 3):
        c, r = rng.uniform(-1, 1, dim), rng.uniform(0.5, 2, dim)
        o = OracleTree(c, r)
        o.upload(*random_tree(rng, dim, 600, max_depth=12))
        bits, depth, flags = o.download()
        dmax = int(depth.max())
        e = o.enumerate()
        assert len(e) == dmax + 1 and e[0] == 1
        for d in range(0, dmax + 2):
            got = o.boxes(d)
            if d == 0:
                exp_flags = np.bitwise_and.reduce(flags)
                assert got.shape == (1, 2 * dim + 1)
                assert np.array_equal(got[0, :dim], c) and np.array_equal(got[0, dim:2 * dim], r)
                assert got[0, -1] == exp_flags
                continue
            groups, gid = depth_groups(bits, depth, d)
            if d <= dmax:
                assert e[d] == len(groups)
            assert got.shape == (len(groups), 2 * dim + 1), d
            for g, a in enumerate(groups):
                lo, hi = bisect_box(a, d, c, r)
                # the bisection rounds at every midpoint; a wrong bisection count is off by 2x
                assert np.allclose(got[g, :dim], 0.5 * (lo + hi), rtol=0, atol=1e-12 * np.abs(r).max())
                assert np.allclose(got[g, dim:2 * dim], 0.5 * (hi - lo), rtol=1e-12, atol=0)
                assert got[g, -1] == np.bitwise_and.reduce(flags[gid == g])
        # the unsubdividable depth: the parents unsubdivide would merge, as many as it leaves
        # behind when every leaf is flagged and their geometry that of the merged boxes
        u = o.boxes(254)
        o.setFlags(0x08, -1)
        o.unsubdivide(0x08)
        b2, d2, _ = o.download()
        merged = {(int(b) >> 0, int(dd)) for b, dd in zip(b2, d2)}
        assert len(u) <= len(b2)
        for row in u:
            hit = False
            for b, dd in merged:
                lo, hi = bisect_box(b, dd, c, r)
                if np.allclose(row[:dim], 0.5 * (lo + hi), atol=1e-12 * np.abs(r).max(), rtol=0) and \
                        np.allclose(row[dim:2 * dim], 0.5 * (hi - lo), rtol=1e-12, atol=0):
                    hit = True
                    break
            assert hit


def hh_energy(x):
    q1, q2, p1, p2 = x[..., 0], x[..., 1], x[..., 2], x[..., 3]
    return 0.5 * (p1 ** 2 + p2 ** 2) + 0.5 * (q1 ** 2 + q2 ** 2) + q1 ** 2 * q2 - q2 ** 3 / 3.0


def leaf_boxes(o, c, r):
    """Centre and radius of every leaf, from the bit pattern (independent of the oracle)."""
    bits, depth, _ = o.download()
    out = []
    for b, d in zip(bits.tolist(), depth.tolist()):
        lo, hi = np.array(c) - np.array(r), np.array(c) + np.array(r)
        for lvl in range(d):  # default scheme: coordinate lvl % 4, bit d-1-lvl of the pattern
            j = lvl % 4
            mid = 0.5 * (lo[j] + hi[j])
            if (b >> (d - 1 - lvl)) & 1:
                lo[j] = mid
            else:
                hi[j] = mid
        out.append((lo, hi))
    return out


def test_energy_band_predicate_contains_the_energy_range():
    """The band predicate's interval bound contains H at every sampled point of every box
    (soundness), and on small boxes its width stays within a few times the sampled range."""
    rng = np.random.default_rng(21)
    c, r = [0.0, 0.0, 0.0, 0.0], [0.6, 0.6, 0.5, 0.5]
    o = OracleTree(c, r)
    bits, depth, flags = random_tree(rng, 4, 600, max_depth=16)
    o.upload(bits, depth, flags)
    boxes = leaf_boxes(o, c, r)
    for emin, emax in [(0.11, 0.12), (0.0, 0.01), (-0.05, 0.0), (0.2, 0.3)]:
        st = o.energy_band_stencil(emin, emax)
        for (lo, hi), s in zip(boxes, st):
            x = lo + (hi - lo) * rng.random((400, 4))
            corners = np.array([[lo[j] if (m >> j) & 1 else hi[j] for j in range(4)] for m in range(16)])
            h = hh_energy(np.vstack([x, corners]))
            if np.any((h >= emin) & (h <= emax)):
                assert s == 1, "a box containing band points was excluded"
    # tightness on small boxes: the predicate keeps few boxes whose sampled range misses the band
    o2 = OracleTree(c, r)
    deep = random_tree(rng, 4, 3000, max_depth=24)
    o2.upload(*deep)
    boxes = leaf_boxes(o2, c, r)
    st = o2.energy_band_stencil(0.11, 0.12)
    far = 0
    for (lo, hi), s in zip(boxes, st):
        x = lo + (hi - lo) * rng.random((200, 4))
        h = hh_energy(x)
        width = h.max() - h.min()
        if s and (h.min() > 0.12 + 4 * width + 1e-12 or h.max() < 0.11 - 4 * width - 1e-12):
            far += 1
    assert far == 0


def test_subspace_iteration_cycle_walk_closed_form():
    """Lazy random walk on a 20-cycle: eigenvalues 0.4 + 0.6 cos(2 pi k / 20); the four leading
    are 1, lambda_1 (double) and lambda_2."""
    n, p = 20, 0.3
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i, 1 - 2 * p), ((i + 1) % n, p), ((i - 1) % n, p)):
            rows.append(i); cols.append(j); vals.append(v)
    o = np.lexsort((cols, rows))
    m = OCoo.from_arrays(np.array(rows)[o], np.array(cols)[o], np.array(vals)[o], n, n)
    H, _, it, _ = m.subspace_iteration(4, 400)
    assert it == 400
    ev = np.sort(np.linalg.eigvals(H).real)[::-1]
    lam = 0.4 + 0.6 * np.cos(2 * np.pi * np.arange(3) / n)
    assert np.allclose(ev, [lam[0], lam[1], lam[1], lam[2]], rtol=0, atol=1e-10)


def test_subspace_iteration_matches_numpy_eig():
    rng = np.random.default_rng(22)
    n = 60
    A = rng.random((n, n)) * (rng.random((n, n)) < 0.2)
    # four weakly coupled blocks give four real leading eigenvalues close to 1 (almost-invariant sets)
    for b in range(4):
        A[b * 15:(b + 1) * 15, b * 15:(b + 1) * 15] += 2.0 * rng.random((15, 15))
    P = A / A.sum(axis=1, keepdims=True)
    r, c = np.nonzero(P)
    H, X, it, _ = OCoo.from_arrays(r, c, P[r, c], n, n).subspace_iteration(4, 300, want_basis=True)
    ev = np.sort(np.linalg.eigvals(H).real)[::-1]
    w = np.linalg.eigvals(P.T)
    ref = np.sort(w[np.argsort(-np.abs(w))[:4]].real)[::-1]
    assert np.allclose(ev, ref, rtol=0, atol=1e-10)
    assert np.allclose(X @ X.T, np.eye(4), atol=1e-12)


def test_subspace_iteration_stopping_rule():
    """With a tolerance the iteration stops at the first basis update k whose residual
    rho = ||X P - H^T X||_F / ||H||_F (rows of X the basis vectors) is at most tol: rho recomputed
    with numpy from the returned basis and Ritz matrix agrees, the fixed-count run to k - 1 has a
    residual above tol, and the Ritz values are those of numpy's eig to the residual's order."""
    rng = np.random.default_rng(23)
    n = 80
    A = rng.random((n, n)) * (rng.random((n, n)) < 0.2)
    for b in range(4):
        A[b * 20:(b + 1) * 20, b * 20:(b + 1) * 20] += 2.0 * rng.random((20, 20))
    P = A / A.sum(axis=1, keepdims=True)
    r, c = np.nonzero(P)
    m = OCoo.from_arrays(r, c, P[r, c], n, n)
    for tol in (1e-6, 1e-10):
        H, X, k, rho = m.subspace_iteration(4, 10000, want_basis=True, tol=tol)
        assert 0 < k < 10000 and rho <= tol
        R = X @ P - H.T @ X
        assert abs(np.linalg.norm(R) / np.linalg.norm(H) - rho) <= 1e-13
        _, _, k1, rho1 = m.subspace_iteration(4, k - 1)
        assert k1 == k - 1 and rho1 > tol
        _, _, k2, rho2 = m.subspace_iteration(4, k)
        assert k2 == k and rho2 == rho
        w = np.linalg.eigvals(P.T)
        ref = np.sort(w[np.argsort(-np.abs(w))[:4]].real)[::-1]
        assert np.allclose(np.sort(np.linalg.eigvals(H).real)[::-1], ref, rtol=0, atol=100 * tol)


def test_subspace_iteration_leading_columns_rule():
    """The residual of the leading `lead` columns: with 8 vectors and lead = 4 the iteration stops
    once the span of the first 4 basis vectors is invariant. Cholesky QR keeps leading spans
    nested, so those 4 vectors are the 4-vector iteration's basis (same start columns) to
    rounding; their span is the dominant 4-dimensional left invariant subspace of P (principal
    angles against numpy's eigenvectors); the eigenvalues of the leading 4 x 4 block of H are
    numpy's 4 leading ones; rho recomputed with numpy from X_L and H_LL agrees."""
    rng = np.random.default_rng(24)
    n = 90
    A = rng.random((n, n)) * (rng.random((n, n)) < 0.2)
    for b in range(4):
        A[b * 20:(b + 1) * 20, b * 20:(b + 1) * 20] += 3.0 * rng.random((20, 20))
    P = A / A.sum(axis=1, keepdims=True)
    r, c = np.nonzero(P)
    m = OCoo.from_arrays(r, c, P[r, c], n, n)
    tol = 1e-10
    H, X, k, rho = m.subspace_iteration(8, 10000, want_basis=True, tol=tol, lead=4)
    assert 0 < k < 10000 and rho <= tol
    XL, HL = X[:4], H[:4, :4]
    assert abs(np.linalg.norm(XL @ P - HL.T @ XL) / np.linalg.norm(HL) - rho) <= 1e-13
    _, _, k1, rho1 = m.subspace_iteration(8, k - 1, lead=4)
    assert rho1 > tol
    # nested spans: the same first 4 vectors as the 4-vector run after k updates
    _, X4, _, _ = m.subspace_iteration(4, k, want_basis=True)
    assert np.allclose(XL, X4, rtol=0, atol=1e-9)
    # the dominant left invariant subspace
    w, V = np.linalg.eig(P.T)
    top = np.argsort(-np.abs(w))[:4]
    Q, _ = np.linalg.qr(np.real_if_close(V[:, top]).real)
    s = np.linalg.svd(Q.T @ XL.T, compute_uv=False)
    assert np.all(s >= 1 - 1e-8)  # cosines of the principal angles
    ref = np.sort(w[top].real)[::-1]
    assert np.allclose(np.sort(np.linalg.eigvals(HL).real)[::-1], ref, rtol=0, atol=1e-8)
    # the whole 8-vector block is a different (stricter) rule
    _, _, k8, rho8 = m.subspace_iteration(8, k)
    assert k8 == k and rho8 > rho


# ------------------------------------------------------------------ HIP library C ABI
def header_symbols():
    text = open(os.path.join(ROOT, "include", "b12x_c_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(b12x_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("flavour", ["parity", "fast"])
def test_library_exports_every_declared_symbol(flavour):
    path = gx.library_path(flavour)
    assert os.path.exists(path), f"{path} not built"
    lib = ctypes.CDLL(path)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # every declared entry point is bound by the Python mirror as well
    from b12_amd.api import SIGNATURES
    assert set(SIGNATURES) | {"b12x_last_error"} >= set(syms)


def test_cpp_facade_builds():
    for name in ("doc_example", "continuation", "api_surface"):
        exe = os.path.join(ROOT, "tests", "cpp", name)
        assert os.path.exists(exe), f"façade example {name} not built"


# ------------------------------------------------------------------ insertion and deleteDepth
def point_pattern(p, d, c, r):
    """Pattern of the depth-d box holding p (default scheme), by bisection (independent of the
    oracle); None outside the domain."""
    dim = len(c)
    lo, hi = np.array(c, float) - np.array(r, float), np.array(c, float) + np.array(r, float)
    if np.any(p < lo) or np.any(p >= hi):
        return None
    a = 0
    for lvl in range(d):
        j = lvl % dim
        mid = 0.5 * (lo[j] + hi[j])
        bit = 1 if p[j] >= mid else 0
        a = (a << 1) | bit
        if bit:
            lo[j] = mid
        else:
            hi[j] = mid
    return a


def test_delete_depth_against_truncation():
    """deleteDepth(d): every leaf truncated to d-1, equal boxes merged with the AND of their
    flags, count(d) returned; deleteDepth(0) leaves the surrounding box alone."""
    rng = np.random.default_rng(81)
    for trial in range(3):
        o = general_tree(rng, 2, 500, 10)
        bits, depth, flags = o.download()
        for d in (3, 6, 9):
            o.upload(bits, depth, flags)
            exp = {}
            order = []
            for b, dd, f in zip(bits.tolist(), depth.tolist(), flags.tolist()):
                key = (b >> (dd - (d - 1)), d - 1) if dd > d - 1 else (b, dd)
                if key not in exp:
                    exp[key] = 0xFF
                    order.append(key)
                exp[key] &= f
            n_d = o.count(d)
            assert o.deleteDepth(d) == n_d
            b2, d2, f2 = o.download()
            assert list(zip(b2.tolist(), d2.tolist())) == order
            assert f2.tolist() == [exp[k] for k in order]
    o.upload(bits, depth, flags)
    assert o.deleteDepth(0) == 1
    assert [a.tolist() for a in o.download()] == [[0], [0], [0]]
    assert o.deleteDepth(254) == 0 and o.deleteDepth(255) == 0


def test_insert_at_a_depth_of_equal_boxes():
    """Leaves all at depth d: inserting points at d adds the boxes of the points no leaf holds
    (isToInsert), with flagsForInserted; the leaves the points hit get flagsForHit."""
    rng = np.random.default_rng(82)
    c, r, d = [0.0, 0.0], [1.0, 1.0], 8
    o = OracleTree(c, r)
    pts0 = rng.uniform(-1, 1, (300, 2))
    o.insertByPoints(pts0, d, 0x02)  # a first cover at depth d (the root's run takes one box)
    bits, depth, flags = o.download()
    assert np.all(depth == d)
    have = set(bits.tolist())
    pts = rng.uniform(-1.2, 1.2, (400, 2))
    res = o.insertByPoints(pts, d, 0x20, 0x40)
    b2, d2, f2 = o.download()
    exp_new, hit = set(), set()
    for i, p in enumerate(pts):
        a = point_pattern(p, d, c, r)
        if a is None:
            assert res[i] == 0
        elif a in have:
            assert res[i] == 0
            hit.add(a)
        else:
            assert res[i] == 1
            exp_new.add(a)
    assert set(b2.tolist()) == have | exp_new and np.all(d2 == d)
    for b, f in zip(b2.tolist(), f2.tolist()):
        old = dict(zip(bits.tolist(), flags.tolist())).get(b)
        if b in exp_new:
            assert f == 0x20
        elif b in hit:
            assert f == old | 0x40
        else:
            assert f == old


def test_insert_into_the_surrounding_box_merges_its_run():
    """The surrounding box sorts between the halves (its comparable key ends its lower half), so
    after inserting boxes of both halves it joins the last lower box and the first upper box in
    one run on a path, which becomes the box of largest pattern (ImplicitBoxTree.hpp:1860-1878):
    the reference's result, kept as it is. Points of one half only keep every box."""
    c, r, d = [0.0, 0.0], [1.0, 1.0], 4
    o = OracleTree(c, r)
    o.deleteDepth(0)
    res = o.insertByPoints([[-0.5, 0.1], [0.3, 0.3], [0.31, 0.29]], d, 0x02)
    assert res.tolist() == [1, 1, 1]
    b, dd, f = o.download()
    up = point_pattern(np.array([0.3, 0.3]), d, c, r)
    assert b.tolist() == [up] and dd.tolist() == [d] and f.tolist() == [0x02]
    o.deleteDepth(0)
    low = [[-0.5, 0.1], [-0.9, -0.2], [-0.1, 0.7]]
    o.insertByPoints(low, d, 0x02)
    assert sorted(o.download()[0].tolist()) == sorted({point_pattern(np.array(p), d, c, r) for p in low})


def test_insert_at_leaf_depth_into_holes():
    """At leaf depth a point outside every leaf becomes a box one level below its deepest common
    ancestor with the leaves on either side: it holds the point, overlaps no leaf, and its parent
    overlaps a neighbour."""
    rng = np.random.default_rng(83)
    c, r = [0.0, 0.0], [1.0, 1.0]
    o = OracleTree(c, r)
    o.upload(*random_tree(rng, 2, 200, max_depth=9, flag_prob=0.0))
    for _ in range(20):
        bits, depth, _ = o.download()
        p = rng.uniform(-1, 1, 2)
        if o.searchByPoints(p[None, :])[0] != 0xFFFFFFFF:
            continue
        res = o.insertByPoints(p[None, :], 255, 0x02)
        assert res[0] == 1
        b2, d2, f2 = o.download()
        new = [(b, dd) for b, dd in zip(b2.tolist(), d2.tolist()) if (b, dd) not in set(zip(bits.tolist(), depth.tolist()))]
        assert len(new) == 1 and len(b2) == len(bits) + 1
        nb, nd = new[0]
        assert point_pattern(p, nd, c, r) == nb
        def overlap(b1, d1, b2_, d2_):
            k = min(d1, d2_)
            return (b1 >> (d1 - k)) == (b2_ >> (d2_ - k))
        assert not any(overlap(nb, nd, b, dd) for b, dd in zip(bits.tolist(), depth.tolist()))
        assert any(overlap(nb >> 1, nd - 1, b, dd) for b, dd in zip(bits.tolist(), depth.tolist()))


def henon_py(x, a=1.4, b=0.3):
    return np.array([1.0 - a * x[0] * x[0] + x[1], b * x[0]])


def sampled_closure(start, d, c, r, n_grid, f):
    """Depth-d boxes reachable from the boxes `start` when every box sends the points of an
    n_grid^dim inner grid (Grids.hpp:105-120) through f: breadth-first search over patterns,
    independent of the oracle."""
    dim = len(c)
    g = np.array([[((i // n_grid ** j) % n_grid * 2.0 + 1.0 - n_grid) / n_grid for j in range(dim)]
                  for i in range(n_grid ** dim)])
    seen, todo = set(start), list(start)
    while todo:
        b = todo.pop()
        lo, hi = bisect_box(b, d, c, r)
        cen, rad = 0.5 * (lo + hi), 0.5 * (hi - lo)
        for x in cen + rad * g:
            a = point_pattern(f(x), d, c, r)
            if a is not None and a not in seen:
                seen.add(a)
                todo.append(a)
    return seen


def test_continuation_is_the_sampled_closure_of_the_start_box():
    """The continuation (continuation.m) from the Hénon map's fixed point ends with exactly the
    depth-d boxes reachable from the fixed point's box under the sampled map, with no flags left;
    the box counts grow until the last step repeats the one before it."""
    a, b = 1.4, 0.3
    x = (-(1 - b) + np.sqrt((1 - b) ** 2 + 4 * a)) / (2 * a)
    p = np.array([[x, b * x]])
    c, r = [0.0, 0.0], [3.0, 3.0]
    o = OracleTree(c, r)
    for d in (6, 10, 13):
        counts = gx.continuation(o, gx.InnerGrid(4), gx.Henon(a, b, 1), [d], p)[0]
        assert counts[0] == 1 and counts[-1] == counts[-2] and all(u < v for u, v in zip(counts[:-2], counts[1:-1]))
        bits, depth, flags = o.download()
        assert np.all(depth == d) and np.all(flags == 0)  # the last step turned INS into EXPD, then cleared it
        exp = sampled_closure([point_pattern(p[0], d, c, r)], d, c, r, 4, henon_py)
        assert set(bits.tolist()) == exp and len(bits) == len(exp) == counts[-1]


# ------------------------------------------------------------------ strong components
def canonical_labels(labels):
    """Relabel a partition with the largest vertex index of each part."""
    labels = np.asarray(labels)
    mx = {}
    for v, l in enumerate(labels.tolist()):
        mx[l] = max(mx.get(l, -1), v)
    return np.array([mx[l] for l in labels.tolist()], np.uint32)


def scipy_scc(rows, cols, n):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    keep = (rows < n) & (cols < n)
    g = coo_matrix((np.ones(int(keep.sum())), (rows[keep], cols[keep])), shape=(n, n)).tocsr()
    return canonical_labels(connected_components(g, directed=True, connection="strong")[1])


def test_strong_components_closed_forms():
    from oracle_api import OCoo
    n = 7
    cyc = OCoo.from_arrays(np.arange(n), (np.arange(n) + 1) % n, np.ones(n), n, n)
    lab, nc, nb = cyc.scc()
    assert lab.tolist() == [n - 1] * n and (nc, nb) == (1, 1)
    path = OCoo.from_arrays(np.arange(n - 1), np.arange(1, n), np.ones(n - 1), n, n)  # a DAG
    lab, nc, nb = path.scc()
    assert lab.tolist() == list(range(n)) and (nc, nb) == (n, 0)
    # two 2-cycles joined by one edge, a self loop, an isolated vertex and an entry outside the
    # vertex range (the column of points outside the cover)
    r = np.array([0, 1, 2, 3, 1, 4, 5])
    c = np.array([1, 0, 3, 2, 2, 4, 6])
    lab, nc, nb = OCoo.from_arrays(r, c, np.ones(7), 6, 7).scc()
    assert lab.tolist() == [1, 1, 3, 3, 4, 5] and (nc, nb) == (4, 2)


def test_strong_components_brute_force_and_scipy():
    """Against reachability by repeated squaring on small graphs and scipy's strong components
    (csgraph.connected_components) on larger ones, including a transition matrix's graph."""
    from oracle_api import OCoo
    rng = np.random.default_rng(1201)
    for trial in range(30):
        n = int(rng.integers(1, 25))
        m = int(rng.integers(0, 3 * n))
        r, c = rng.integers(0, n, m), rng.integers(0, n, m)
        pairs = sorted(set(zip(r.tolist(), c.tolist())))
        r = np.array([p[0] for p in pairs], np.uint32)
        c = np.array([p[1] for p in pairs], np.uint32)
        lab, _, _ = OCoo.from_arrays(r, c, np.ones(len(r)), n, n).scc()
        R = np.eye(n, dtype=bool)
        R[r, c] = True
        for _ in range(6):
            R = R | ((R.astype(np.int64) @ R.astype(np.int64)) > 0)
        same = R & R.T
        exp = np.array([max(np.nonzero(same[v])[0]) for v in range(n)], np.uint32)
        assert np.array_equal(lab, exp)
    for n, m in ((2000, 3000), (5000, 20000), (3000, 2500)):
        r, c = rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32)
        lab, _, _ = OCoo.from_arrays(r, c, np.ones(m), n, n).scc()
        assert np.array_equal(lab, scipy_scc(r, c, n))
    o = OracleTree([0.0, 0.0], [1.5, 0.5])
    o.setFlags(gx.SD, 0)
    o.subdivision_steps(gx.InnerGrid(4), gx.Henon(), 12)
    m = o.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon())
    r, c, _ = m.download()
    lab, nc, nb = m.scc()
    assert np.array_equal(lab, scipy_scc(r, c, m.nRows)) and nc == len(set(lab.tolist()))


def test_fixed_boxes_against_a_restatement_and_the_fixed_points():
    """fixedBoxes.m: per step every box splits along the scheme and a child is kept when one of
    its inner-grid points is mapped into the child itself. Restated here on bit patterns with
    bisection (independent of the oracle); every kept box lies within two box radii of one of the
    Hénon map's fixed points (sampling can lose a fixed point's boxes: a box is kept only when one
    of its 16 points lands in it)."""
    a, b = 1.4, 0.3
    c, r = [0.0, 0.0], [3.0, 3.0]
    g = np.array([[((i // 4 ** j) % 4 * 2.0 + 1.0 - 4) / 4 for j in range(2)] for i in range(16)])
    steps = 24
    boxes = [0]
    for d in range(1, steps + 1):
        kept = []
        for p in boxes:
            for child in (2 * p, 2 * p + 1):
                lo, hi = bisect_box(child, d, c, r)
                cen, rad = 0.5 * (lo + hi), 0.5 * (hi - lo)
                if any(point_pattern(henon_py(x, a, b), d, c, r) == child for x in cen + rad * g):
                    kept.append(child)
        boxes = kept if kept else [0]
        if not kept:
            break
    o = OracleTree(c, r)
    B = gx.fixedBoxes(o, gx.InnerGrid(4), gx.Henon(a, b, 1), steps)
    bits, depth, _ = o.download()
    assert sorted(bits.tolist()) == sorted(boxes) and np.all(depth == steps)
    fps = [np.array([x, b * x]) for x in ((-(1 - b) + s * np.sqrt((1 - b) ** 2 + 4 * a)) / (2 * a) for s in (1, -1))]
    dist = np.array([[np.max(np.abs(fp - row[:2]) / row[2:4]) for fp in fps] for row in B])
    assert len(B) >= 1 and np.all(dist.min(axis=1) <= 2.0)


def test_subdivision_by_nim_cover_inside_the_plain_subdivision():
    """subdivisionByNim.m keeps the leaves whose invariant measure is at least 1e-9 / n. A box of
    positive measure is hit by an image of a kept box, so by induction over the depths the cover
    lies inside the plain subdivision's cover (subdivision.m) at every depth; the measure of the
    kept boxes carries almost all of the mass."""
    grid, mp = gx.InnerGrid(4), gx.Henon()
    o, p = OracleTree([0.0, 0.0], [1.5, 0.5]), OracleTree([0.0, 0.0], [1.5, 0.5])
    p.setFlags(gx.SD, gx.LEAF)
    o.setFlags(gx.SD, gx.LEAF)
    for d in range(1, 17):
        got = gx.subdivisionByNim(o, grid, mp, 1)[0]
        p.subdivision_steps(grid, mp, 1)
        kept = set(zip(*[x.tolist() for x in o.download()[:2]]))
        plain = set(zip(*[x.tolist() for x in p.download()[:2]]))
        assert kept and kept <= plain and got[2] == len(kept), d


def test_json_stream_helper_writes_raw_uint8_bytes():
    """tests/cpp/json_stream (the tree file as iostreams write it, used by the file tests) puts
    one raw byte per uint8 element and reads one character per token."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "json_stream")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/json_stream not built")
    exp = ('{\n  "B12_VERSION_MAJOR": 3,\n  "B12_VERSION_MINOR": 1,\n  "center": [\n    0.10000000000000001\n  ],'
           '\n  "radius": [\n    1\n  ],\n  "sdScheme": [\n    \x00\n  ],\n  "nLeaves": 2,'
           '\n  "bitPatternVector": [\n    0,\n    1\n  ],\n  "depthVector": [\n    \x14,\n    \x14\n  ],'
           '\n  "flagsVector": [\n    \x08,\n    \t\n  ],\n  "nHostBufferBytes": 4294967296\n}').encode()
    out = subprocess.run([exe, "write"], input=b"1 0.1 1 1 0 2 0 1 20 20 8 9 4294967296", capture_output=True,
                         check=True).stdout
    assert out == exp
    back = subprocess.run([exe, "read", "flagsVector"], input=out, capture_output=True, check=True).stdout
    assert back.split() == [b"8"]  # the tab (SD|HIT) is lost as whitespace
    back = subprocess.run([exe, "read", "depthVector"], input=out, capture_output=True, check=True).stdout
    assert back.split() == [b"20", b"20"]


def _substochastic(rng, n, density, blocks=4):
    """A random sparse row-substochastic matrix with `blocks` weakly coupled diagonal blocks."""
    A = rng.random((n, n)) * (rng.random((n, n)) < density)
    s = n // blocks
    for b in range(blocks):
        A[b * s:(b + 1) * s, b * s:(b + 1) * s] += 3.0 * rng.random((s, s)) * (rng.random((s, s)) < 4 * density)
    P = 0.97 * A / np.maximum(A.sum(axis=1, keepdims=True), 1e-300)
    r, c = np.nonzero(P)
    return P, r, c


def test_krylov_basis_operations():
    """The oracle's Krylov basis: start normalises v0; each expansion gives the Arnoldi relation
    V_p^T P ... in row-vector form, w = v_p P = V_{p+1} h + beta v_{p+1}, with V orthonormal to
    rounding; rotate forms V Q and moves the residual vector. Checked with numpy from the vectors
    the basis returns (not by restating its sums)."""
    rng = np.random.default_rng(5)
    n, ncv = 300, 12
    P, r, c = _substochastic(rng, n, 0.05)
    m = OCoo.from_arrays(r, c, P[r, c], n, n)
    K = m.krylov(ncv)
    v0 = rng.standard_normal(n)
    nrm = K.start(v0)
    assert np.isclose(nrm, np.linalg.norm(v0), rtol=1e-14)
    assert np.allclose(K.vector(0), v0 / np.linalg.norm(v0), rtol=0, atol=1e-15)
    for p in range(ncv):
        vp = K.vector(p)
        h, beta = K.expand(p)
        V = np.stack([K.vector(q) for q in range(p + 2)])
        assert np.allclose(V @ V.T, np.eye(p + 2), atol=1e-13)
        w = vp @ P
        assert np.allclose(w, h @ V[:p + 1] + beta * V[p + 1], rtol=0, atol=1e-14)
        assert np.allclose(h, V[:p + 1] @ w, rtol=0, atol=1e-14)
    Q, _ = np.linalg.qr(rng.standard_normal((ncv, 5)))
    V = np.stack([K.vector(q) for q in range(ncv + 1)])
    K.rotate(ncv, 5, Q)
    assert np.allclose(np.stack([K.vector(q) for q in range(5)]), Q.T @ V[:ncv], rtol=0, atol=1e-14)
    assert np.array_equal(K.vector(5), V[ncv])
    # the default start: the subspace iteration's second start column, normalised
    K.start(None)
    u = np.array([((_splitmix64(8 * i + 1) >> 11) * 2.0 ** -53) * 2.0 - 1.0 for i in range(n)])
    assert np.allclose(K.vector(0), u / np.linalg.norm(u), rtol=0, atol=1e-15)


def _splitmix64(x):
    M = (1 << 64) - 1
    x = (x + 0x9e3779b97f4a7c15) & M
    x = ((x ^ (x >> 30)) * 0xbf58476d1ce4e5b9) & M
    x = ((x ^ (x >> 27)) * 0x94d049bb133111eb) & M
    return x ^ (x >> 31)


def test_eigs_directed_lazy_cycle_closed_form():
    """Lazy walk around a directed 60-cycle, P = (1 - a) I + a C: eigenvalues (1 - a) + a w^k,
    w = exp(2 pi i / 60), all simple. The leading by modulus are 1, then the k = +-1 pair, then
    the k = +-2 pair: nev = 4 cuts that pair, so both of it are returned (5 values)."""
    n, a = 60, 0.3
    rows = np.concatenate([np.arange(n), np.arange(n)])
    cols = np.concatenate([np.arange(n), (np.arange(n) + 1) % n])
    vals = np.concatenate([np.full(n, 1 - a), np.full(n, a)])
    o = np.lexsort((cols, rows))
    m = OCoo.from_arrays(rows[o], cols[o], vals[o], n, n)
    res = gx.eigs(m, nev=4, ncv=30, tol=1e-12)
    assert res["converged"]
    lam = (1 - a) + a * np.exp(2j * np.pi * np.array([0, 1, -1, 2, -2]) / n)
    got = res["eigenvalues"]
    assert len(got) == 5
    for z in lam:
        assert np.min(np.abs(got - z)) < 1e-11
    assert res["residual"] <= 1e-12 and np.all(res["pair_residuals"] <= 1e-10)


def test_eigs_matches_scipy_on_random_substochastic_matrices():
    """The four (or five, with a cut pair) leading eigenvalues of random sparse substochastic
    matrices with almost-invariant blocks equal scipy's eigs (ARPACK) to 1e-9, and the returned
    Ritz vectors are left eigenvectors to the stated residual."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    for seed, n in ((11, 400), (12, 1500)):
        rng = np.random.default_rng(seed)
        P, r, c = _substochastic(rng, n, 8.0 / n)
        m = OCoo.from_arrays(r, c, P[r, c], n, n)
        res = gx.eigs(m, nev=4, tol=1e-11, want_vectors=True)
        assert res["converged"]
        w = sla.eigs(sp.csr_matrix(P).T.tocsr(), k=8, which="LM", tol=1e-14, maxiter=100000)[0]
        w = w[np.argsort(-np.abs(w))]
        for z in res["eigenvalues"]:
            assert np.min(np.abs(w - z)) <= 1e-9 * abs(z)
        for z, v in zip(res["eigenvalues"], res["vectors"]):
            assert np.linalg.norm(v @ P - z * v) <= 1e-10 * abs(z) * np.linalg.norm(v)
