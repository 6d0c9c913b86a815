"""Strong components of a matrix's graph and the chain recurrent set step (chainRecurrentSet.m)
on the GPU against the oracle: labels (the largest vertex index of each component) must be
identical, and the trees after each step identical."""
import numpy as np
import pytest

import b12_amd as gx
from helpers import same_tree
from oracle_api import OCoo, OracleTree

pytestmark = pytest.mark.gpu


def both_scc(rows, cols, n, n_cols=None):
    n_cols = n if n_cols is None else n_cols
    vals = np.ones(len(rows))
    g = gx.CooMatrix.from_arrays(rows, cols, vals, n, n_cols)
    o = OCoo.from_arrays(rows, cols, vals, n, n_cols)
    return g.scc(), o.scc()


def unique_edges(r, c):
    pairs = sorted(set(zip(np.asarray(r).tolist(), np.asarray(c).tolist())))
    return (np.array([p[0] for p in pairs], np.uint32), np.array([p[1] for p in pairs], np.uint32))


@pytest.mark.parametrize("n,m,seed", [(1, 0, 1), (1, 1, 2), (50, 0, 3), (200, 150, 4), (2000, 2500, 5),
                                      (5000, 20000, 6), (100000, 150000, 7), (100000, 400000, 8)])
def test_scc_random_graphs_match_oracle(n, m, seed):
    print(f"[scc] random n={n} m={m}", flush=True)
    rng = np.random.default_rng(1300 + seed)
    r, c = unique_edges(rng.integers(0, n, m), rng.integers(0, n, m))
    (lg, ncg, nbg), (lo, nco, nbo) = both_scc(r, c, n)
    assert np.array_equal(lg, lo) and (ncg, nbg) == (nco, nbo)


def test_scc_chains_cycles_and_outside_column():
    """Long paths in both index directions (trimmed pass by pass), one long cycle, a cycle with
    chords, self loops, and entries in the extra column of points outside the cover."""
    L = 3000
    cases = [
        (np.arange(L - 1), np.arange(1, L), L),           # increasing path
        (np.arange(1, L), np.arange(L - 1), L),           # decreasing path
        (np.arange(L), (np.arange(L) + 1) % L, L),        # one cycle
        (np.arange(L), (np.arange(L) * 7 + 3) % L, L),    # a permutation: several cycles
    ]
    for r, c, n in cases:
        print(f"[scc] structured n={n}", flush=True)
        r, c = unique_edges(r, c)
        (lg, ncg, nbg), (lo, nco, nbo) = both_scc(r, c, n)
        assert np.array_equal(lg, lo) and (ncg, nbg) == (nco, nbo)
    # two cycles joined by a path, self loops, an isolated vertex, the extra column
    r = np.array([0, 1, 1, 2, 3, 4, 5, 5, 6, 7, 8, 2, 6])
    c = np.array([1, 0, 2, 3, 4, 5, 6, 7, 5, 7, 9, 2, 10])
    r, c = unique_edges(r, c)
    (lg, ncg, nbg), (lo, nco, nbo) = both_scc(r, c, 10, 11)
    assert np.array_equal(lg, lo) and (ncg, nbg) == (nco, nbo)
    assert lg.tolist() == [1, 1, 2, 3, 4, 6, 6, 7, 8, 9]


def test_scc_of_transition_matrices():
    for grid, mp, c, r, steps in ((gx.InnerGrid(4), gx.Henon(), [0.0, 0.0], [1.5, 0.5], 18),
                                  (gx.InnerGrid(2), gx.Lorenz(), [0.0, 0.0, 27.0], [30.0, 30.0, 40.0], 21)):
        print(f"[scc] transition matrix steps={steps}", flush=True)
        g = gx.ImplicitBoxTree(c, r)
        g.setFlags(gx.SD, 0)
        g.subdivision_steps(grid, mp, steps)
        for use_invalid in (False, True):
            m = g.transitionMatrixForLeaves(grid, mp, use_invalid=use_invalid)
            rows, cols, vals = m.download()
            o = OCoo.from_arrays(rows, cols, vals, m.nRows, m.nColumns)
            lg, ncg, nbg = m.scc()
            lo, nco, nbo = o.scc()
            assert np.array_equal(lg, lo) and (ncg, nbg) == (nco, nbo)


@pytest.mark.parametrize("name,grid,mp,c,r,steps", [
    ("henon", gx.InnerGrid(4), gx.Henon(), [0.0, 0.0], [3.0, 3.0], 24),
    ("lorenz", gx.InnerGrid(2), gx.Lorenz(), [0.0, 0.0, 27.0], [30.0, 30.0, 40.0], 21),
])
def test_chain_recurrent_steps_match_oracle(name, grid, mp, c, r, steps):
    print(f"[scc] chain recurrent {name} steps={steps}", flush=True)
    o = OracleTree(c, r)
    exp = gx.chainRecurrentSet(o, grid, mp, steps)
    # fused on the device
    g = gx.ImplicitBoxTree(c, r)
    g.setFlags(gx.SD, gx.LEAF)
    got = g.chain_recurrent_steps(grid, mp, steps)
    assert [tuple(int(v) for v in row) for row in got] == [tuple(e) for e in exp]
    assert same_tree(g.download(), o.download())
    # the same driver as the oracle's, over the GPU tree and its matrices
    h = gx.ImplicitBoxTree(c, r)
    assert gx.chainRecurrentSet(h, grid, mp, steps) == exp
    assert same_tree(h.download(), o.download())


def test_scc_and_power_iteration_in_any_entry_order():
    """Entries uploaded in column-major order, in no order, or with an entry repeated give the
    same components; the power iteration of a shuffled upload equals that of the sorted matrix
    bit for bit (its sums run over each column in ascending row order either way) and the
    oracle's within 1e-12 relative (L1)."""
    rng = np.random.default_rng(1401)
    n, m = 20000, 60000
    r, c = unique_edges(rng.integers(0, n, m), rng.integers(0, n, m))
    lo, nco, nbo = OCoo.from_arrays(r, c, np.ones(len(r)), n, n).scc()
    cm = np.lexsort((r, c))
    sh = rng.permutation(len(r))
    rep = np.concatenate([sh, sh[:500]])
    for order in (cm, sh, rep):
        print(f"[scc] order of {len(order)} entries", flush=True)
        g = gx.CooMatrix.from_arrays(r[order], c[order], np.ones(len(order)), n, n)
        lg, ncg, nbg = g.scc()
        assert np.array_equal(lg, lo) and (ncg, nbg) == (nco, nbo)
    t = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    t.setFlags(gx.SD, 0)
    t.subdivision_steps(gx.InnerGrid(4), gx.Henon(), 14)
    a = t.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), row_major=True)
    rows, cols, vals = a.download()
    ref = a.power_iteration(30)[0]
    sh = rng.permutation(len(rows))
    b = gx.CooMatrix.from_arrays(rows[sh], cols[sh], vals[sh], a.nRows, a.nColumns)
    assert np.array_equal(b.power_iteration(30)[0], ref)
    o = OCoo.from_arrays(rows[sh], cols[sh], vals[sh], a.nRows, a.nColumns)
    mu_o = o.power_iteration(30)[0]
    assert np.array_equal(mu_o, ref)  # bit for bit (the measure's bar: 1e-12 relative)
