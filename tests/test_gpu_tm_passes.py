"""The transition matrix in passes (ImplicitBoxTree.hpp:2941-3053: the reference sizes every pass
from the free device memory and carries the open last row across passes; here a pass ends at a
box boundary). With the pass scratch bounded to a few boxes the matrix is the single-pass one
and the oracle's, bit for bit, for both test-point paths (k <= 64 per-box segments, k > 64 sort),
every flag combination, leaf, general-depth and root sources and a seeded UniformGrid, whose
points depend on their global index."""
import numpy as np
import pytest

import b12_amd as gx
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu


def run(center, radius, grid, mp, steps):
    g, o = gx.ImplicitBoxTree(center, radius), OracleTree(center, radius)
    for t in (g, o):
        t.setFlags(gx.SD, 0)
        t.subdivision_steps(grid, mp, steps)
    return g, o


def same(a, b):
    assert (a.nnz, a.nRows, a.nColumns) == (b.nnz, b.nRows, b.nColumns)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)


def passes_for(boxes_per_pass, n):
    return -(-n // boxes_per_pass)


@pytest.mark.parametrize("use_abs,row_major,use_invalid", [(False, False, False), (True, True, True),
                                                           (False, True, False), (False, False, True)])
def test_passes_match_single_pass_and_oracle(use_abs, row_major, use_invalid):
    print(f"[tm passes] henon abs={use_abs} row_major={row_major} invalid={use_invalid}", flush=True)
    grid, mp = gx.InnerGrid(4), gx.Henon()
    g, o = run([0.0, 0.0], [1.5, 0.5], grid, mp, 12)
    st = np.arange(g.size()) % 3 != 0  # some targets fall outside the cover
    for t in (g, o):
        t.setFlagsByStencil(st, 0x20)
        t.remove(0x20)
    n, k = g.size(), 16
    one = g.transitionMatrixForLeaves(grid, mp, use_abs, row_major, use_invalid)
    assert gx.tm_last_passes() == 1
    ref = o.transitionMatrixForLeaves(grid, mp, use_abs, row_major, use_invalid)
    same(one, ref)
    for boxes in (1, 7, 100, n - 1):
        g.set_tm_pass_bytes(4 * k * boxes)
        m = g.transitionMatrixForLeaves(grid, mp, use_abs, row_major, use_invalid)
        assert gx.tm_last_passes() == passes_for(boxes, n), boxes
        same(m, ref)
    g.set_tm_pass_bytes(1)  # below one box: one box a pass
    same(g.transitionMatrixForLeaves(grid, mp, use_abs, row_major, use_invalid), ref)
    assert gx.tm_last_passes() == n
    g.set_tm_pass_bytes(0)
    same(g.transitionMatrixForLeaves(grid, mp, use_abs, row_major, use_invalid), ref)
    assert gx.tm_last_passes() == 1


def test_passes_more_than_64_points_per_box():
    print("[tm passes] k = 100 (sort path)", flush=True)
    grid, mp = gx.InnerGrid(10), gx.Henon()
    g, o = run([0.0, 0.0], [1.5, 0.5], grid, mp, 9)
    n, k = g.size(), 100
    ref = o.transitionMatrixForLeaves(grid, mp, False, False, True)
    for boxes in (3, 50):
        g.set_tm_pass_bytes(24 * k * boxes)
        m = g.transitionMatrixForLeaves(grid, mp, False, False, True)
        assert gx.tm_last_passes() == passes_for(boxes, n)
        same(m, ref)


def test_passes_general_depth_root_and_uniform_grid():
    print("[tm passes] depth, root, UniformGrid", flush=True)
    grid, mp = gx.InnerGrid(3), gx.Lorenz()
    g, o = run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], grid, mp, 12)
    g.set_tm_pass_bytes(4 * 27 * 5)
    same(g.transitionMatrixForDepth(9, grid, mp), o.transitionMatrixForDepth(9, grid, mp))
    assert gx.tm_last_passes() > 1
    same(g.transitionMatrixForRoot(grid, mp, use_abs=True), o.transitionMatrixForRoot(grid, mp, use_abs=True))
    assert gx.tm_last_passes() == 1
    ug = gx.UniformGrid(48, 7)
    ref = o.transitionMatrixForLeaves(ug, mp)
    for boxes in (1, 13):
        g.set_tm_pass_bytes(4 * 48 * boxes)
        same(g.transitionMatrixForLeaves(ug, mp), ref)
        assert gx.tm_last_passes() == passes_for(boxes, g.size())


def test_passes_then_measure():
    """A matrix built in passes gives the single-pass measure bit for bit."""
    print("[tm passes] measure", flush=True)
    grid, mp = gx.InnerGrid(3), gx.Lorenz()
    g, o = run([0.0, 0.0, 27.0], [30.0, 30.0, 30.0], grid, mp, 16)
    mu1, it1, r1 = g.transitionMatrixForLeaves(grid, mp).power_iteration(3000, 1e-12)
    g.set_tm_pass_bytes(4 * 27 * 1000)
    m = g.transitionMatrixForLeaves(grid, mp)
    assert gx.tm_last_passes() == passes_for(1000, g.size())
    mu2, it2, r2 = m.power_iteration(3000, 1e-12)
    assert it1 == it2 and r1 == r2 and np.array_equal(mu1, mu2)
