"""b12x_power_iteration_noalloc (the entry point without device allocation) called through the C ABI
against b12x_power_iteration on the same matrices: equal measures, iteration counts and residuals
bit for bit, on consecutive calls where a smaller matrix follows a larger one (the scratch keeps
the previous call's contents), with a given start vector, and the errors for matrices beyond the
scratch's 131072 columns or not sorted column-major."""
import numpy as np
import pytest

import b12_amd as gx

pytestmark = pytest.mark.gpu


def cover_matrix(center, radius, grid, mp, depth, **kw):
    t = gx.ImplicitBoxTree(center, radius)
    t.setFlags(gx.SD, 0)
    t.subdivision_steps(grid, mp, depth)
    return t.transitionMatrixForLeaves(grid, mp, **kw)


def same(a, b):
    (ma, ia, ra), (mb, ib, rb) = a, b
    return ia == ib and ra == rb and np.array_equal(ma, mb)


def test_noalloc_power_iteration_matches_the_allocating_form():
    print("[noalloc] lorenz d20 then henon d12", flush=True)
    big = cover_matrix([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], gx.InnerGrid(3), gx.Lorenz(), 20)
    small = cover_matrix([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 12)
    assert big.nRows > small.nRows > 100
    for _ in range(2):
        a = big.power_iteration_noalloc(100000, 1e-12)
        assert a[1] > 100 and a[2] <= 1e-12
        assert same(a, big.power_iteration(100000, 1e-12))
        b = small.power_iteration_noalloc(200)
        assert b[1] == 200
        assert same(b, small.power_iteration(200))
    mu0 = np.random.default_rng(3).random(small.nRows)
    assert same(small.power_iteration_noalloc(37, mu=mu0), small.power_iteration(37, mu=mu0))
    assert same(small.power_iteration_noalloc(0), small.power_iteration(0))


def test_noalloc_power_iteration_matrix_with_the_invalid_column():
    m = cover_matrix([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 10, use_invalid=True)
    assert m.nColumns == m.nRows + 1
    assert same(m.power_iteration_noalloc(300, 1e-13), m.power_iteration(300, 1e-13))


def test_noalloc_power_iteration_rejects_what_it_cannot_hold():
    n = 131073
    rows = np.arange(n, dtype=np.uint32)
    cols = np.roll(rows, 1)
    order = np.lexsort((rows, cols))
    big = gx.CooMatrix.from_arrays(rows[order], cols[order], np.ones(n), n, n)
    with pytest.raises(gx.B12xError, match="exceed its scratch maximum of 131072"):
        big.power_iteration_noalloc(5)
    rm = cover_matrix([0.0, 0.0], [1.5, 0.5], gx.InnerGrid(4), gx.Henon(), 8, row_major=True)
    with pytest.raises(gx.B12xError, match="sorted column-major"):
        rm.power_iteration_noalloc(5)
    rm.sortColumnMajor()
    assert same(rm.power_iteration_noalloc(50), rm.power_iteration(50))
