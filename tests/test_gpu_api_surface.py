"""The drop-in surface beyond the hot path: domain getters and setters, the subdivision scheme's
bookkeeping, areAllLeavesOnDepth, hasSameBoxes, copies, the host buffer size kept in files, and
the by-points / by-search flag forms (ImplicitBoxTree.h:46-108, 141-143, 244, 472-480, 532-569).
The C++ façade's program (tests/cpp/api_surface.cpp) must print what the Python mirror computes,
and the flag forms must leave the oracle's tree."""
import os
import subprocess

import numpy as np
import pytest

import b12_amd as gx
from helpers import same_tree
from oracle_api import OracleTree

pytestmark = pytest.mark.gpu

PTS = np.array([[0.1, 0.05], [-0.7, 0.2], [0.9, -0.3], [5.0, 5.0], [0.12, 0.051]])


def stencil(n, search):
    st = np.zeros(n, np.uint8)
    s = np.asarray(search)
    s = s[(s != gx.INVALID) & (s < n)]
    st[s] = 1
    return st


def fmt(name, v):
    return name + "".join(f" {float(x):.17g}" for x in v)


def test_cpp_facade_api_surface_matches_python_and_oracle(tmp_path):
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_surface")
    assert os.path.exists(exe), "tests/cpp/api_surface not built (run __graft_entry__.build())"
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.splitlines()
    out = [l for l in out if not l.startswith("Warning:")]
    grid, mp = gx.InnerGrid(4), gx.Henon()
    t = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    o = OracleTree([0.0, 0.0], [1.5, 0.5])
    for x in (t, o):
        x.setFlags(gx.SD, 0)
        x.subdivision_steps(grid, mp, 10)
    exp = [fmt("center", t.getCenter()), fmt("radius", t.getRadius()), f"default {int(t.isSdSchemeDefault())}",
           f"alldepth {int(t.areAllLeavesOnDepth(gx.LEAF))} {int(t.areAllLeavesOnDepth(10))} "
           f"{int(t.areAllLeavesOnDepth(254))}"]
    u = t.copy()
    exp.append(f"same {int(t.hasSameBoxes(u))}")
    r = t.setFlagsByPoints(PTS, 0x10)
    assert r == o.setFlagsByStencil(stencil(o.count(-1), o.searchByPoints(PTS, -1)), 0x10, -1)
    exp.append(f"setByPoints {r}")
    r = t.changeFlagsByPoints(PTS, 0x10, 0x20, 6)
    assert r == o.changeFlagsByStencil(stencil(o.count(6), o.searchByPoints(PTS, 6)), 0x10, 0x20, 6)
    exp.append(f"changeByPoints {r}")
    r = t.unsetFlagsByPoints(PTS, 0x20, 6)
    assert r == o.unsetFlagsByStencil(stencil(o.count(6), o.searchByPoints(PTS, 6)), 0x20, 6)
    exp.append(f"unsetByPoints {r}")
    exp.append(f"same {int(t.hasSameBoxes(u))}")
    s = t.searchCompacted(t.makePointIteratorPairOverDepth(7, grid, mp), 7)
    assert np.array_equal(s, o.searchCompacted(o.makePointIteratorPairOverDepth(7, grid, mp), 7))
    r = t.changeFlagsBySearch(s, 0x00, 0x40, 7)
    assert r == o.changeFlagsByStencil(stencil(o.count(7), s), 0x00, 0x40, 7)
    exp.append(f"changeBySearch {r}")
    r = t.unsetFlagsBySearch(s, 0x40, 7)
    assert r == o.unsetFlagsByStencil(stencil(o.count(7), s), 0x40, 7)
    exp.append(f"unsetBySearch {r}")
    assert same_tree(t.download(), o.download())
    m = t.transitionMatrixForDepth(7, grid, mp)
    exp.append(f"tm7 {m.nnz} {m.nRows} {m.nColumns}")
    exp.append(f"hostbytes {t.getNHostBufferBytes()} {t.setNHostBufferBytes(1 << 20)}")
    exp.append(fmt("sd", t.setSdScheme([1, 0, 0])))
    exp.append(fmt("sdcount", t.getSdCount()))
    exp.append(f"default {int(t.isSdSchemeDefault())}")
    exp.append(fmt("radius", t.setRadius([-1.0, 2.0])))
    exp.append(fmt("center", t.setCenter([0.25])))
    exp.append(f"same {int(t.hasSameBoxes(u))}")
    w = t.copy()
    exp.append(f"same {int(w.hasSameBoxes(t))} {w.getNHostBufferBytes()}")
    P = u.transitionMatrixForLeaves(grid, mp)
    Q = u.transitionMatrixForLeaves(grid, mp, row_major=True)
    exp.append(f"coo {int(P.isSortedColumnMajor())} {int(P.isSortedRowMajor())} {int(Q.isSortedRowMajor())} "
               f"{int(P == Q)} {int(Q == P)}")
    R = P.copy().sortRowMajor()
    exp.append(f"sorted {int(R.isSortedRowMajor())} "
               f"{int(all(np.array_equal(x, y) for x, y in zip(R.download(), Q.download())))} {int(R == P)}")
    rr, cc, vv = R.download()
    vv[0] += 1.0
    R2 = gx.CooMatrix.from_arrays(rr, cc, vv, R.nRows, R.nColumns)
    exp.append(f"differ {int(R2 == P)}")
    S = Q.copy()
    S.makeSymmetric()
    exp.append(f"symmetric {S.nnz} {int(S.isSortedRowMajor())}")
    assert out == exp
    assert exp[-4:-1] == ["coo 1 0 1 1 1", "sorted 1 1 1", "differ 0"]
    # the values the reference fixes
    assert exp[2] == "default 1" and exp[3] == "alldepth 1 1 0" and exp[4] == "same 1"
    sd = t.getSdScheme()
    assert len(sd) == 65 and list(sd[:6]) == [1, 0, 0, 1, 0, 0]
    assert list(t.getCenter()) == [0.25, 0.5] and list(t.getRadius()) == [1.5, 0.5]
    # the stored host buffer size goes to the file and comes back; t holds depth-10 leaves, which
    # the reference's reader (and so b12x's) cannot read back (byte 10 is whitespace), so a depth-8
    # tree with the same setting makes the round trip
    path = str(tmp_path / "tree")
    t.save(path)
    assert '"nHostBufferBytes": 1048576' in open(path + ".json", "rb").read().decode("latin-1")
    with pytest.raises(gx.B12xError, match="fewer entries than nLeaves"):
        gx.ImplicitBoxTree.load(path, 2)
    v = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    v.setFlags(gx.SD, 0)
    v.subdivision_steps(grid, mp, 8)
    v.setNHostBufferBytes(1 << 20)
    v.save(path)
    assert gx.ImplicitBoxTree.load(path, 2).getNHostBufferBytes() == 1 << 20


def test_has_same_boxes_and_domain_setters():
    """hasSameBoxes against a changed flag, leaf, centre or scheme; a tree with more leaves whose
    first leaves match counts as the same (the reference compares this tree's leaves only); a
    new scheme changes where points fall while the leaves stay."""
    rng = np.random.default_rng(1501)
    a = gx.ImplicitBoxTree([0.0, 0.0], [1.0, 1.0])
    a.setFlags(gx.SD, 0)
    for _ in range(6):
        a.subdivide(gx.SD)
    b = a.copy()
    assert a.hasSameBoxes(b) and b.hasSameBoxes(a)
    b.setFlagsByStencil(np.eye(1, b.size(), 5, dtype=np.uint8)[0], 0x40)
    assert not a.hasSameBoxes(b)
    c = a.copy()
    c.setCenter([0.0, 1e-300])
    assert not a.hasSameBoxes(c)
    bits, depth, flags = a.download()
    d = gx.ImplicitBoxTree([0.0, 0.0], [1.0, 1.0])
    d.upload(bits[:-1], depth[:-1], flags[:-1])
    assert d.hasSameBoxes(a) and not a.hasSameBoxes(d)
    e = a.copy()
    e.setSdScheme([1, 0])
    assert not a.hasSameBoxes(e)  # schemes differ within depth() = 6
    pts = rng.uniform(-1, 1, (200, 2))
    o = OracleTree([0.0, 0.0], [1.0, 1.0], [1, 0])
    o.upload(bits, depth, flags)
    assert np.array_equal(e.searchByPoints(pts), o.searchByPoints(pts))
    assert e.getSdScheme().tolist() == [[1, 0][i % 2] for i in range(65)]
    e.setSdScheme([])
    assert e.isSdSchemeDefault() and a.hasSameBoxes(e)


def test_coo_sorts_are_stable_and_equality_ignores_order():
    """sortRowMajor / sortColumnMajor reorder the entries stably (repeated entries keep their
    order, as numpy's stable lexsort does); isSorted* accept equal neighbours; == holds between
    orderings of one matrix and fails after one value changes."""
    rng = np.random.default_rng(1601)
    n, m = 3000, 20000
    r = rng.integers(0, n, m).astype(np.uint32)
    c = rng.integers(0, n, m).astype(np.uint32)
    r[:100], c[:100] = r[100:200], c[100:200]  # repeated entries
    v = rng.random(m)
    a = gx.CooMatrix.from_arrays(r, c, v, n, n)
    assert not a.isSortedRowMajor() and not a.isSortedColumnMajor()
    rm = np.lexsort((c, r))
    cm = np.lexsort((r, c))
    b = a.copy().sortRowMajor()
    assert b.isSortedRowMajor()
    for x, y in zip(b.download(), (r[rm], c[rm], v[rm])):
        assert np.array_equal(x, y)
    d = a.copy().sortColumnMajor()
    assert d.isSortedColumnMajor()
    for x, y in zip(d.download(), (r[cm], c[cm], v[cm])):
        assert np.array_equal(x, y)
    assert a == b and b == d and d == a
    v2 = v.copy()
    v2[7] += 0.5
    assert not (gx.CooMatrix.from_arrays(r, c, v2, n, n) == a)
    assert not (gx.CooMatrix.from_arrays(r, c, v, n, n + 1) == a)
    # the power iteration of a matrix sorted in place equals that of its row-major source
    t = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    t.setFlags(gx.SD, 0)
    t.subdivision_steps(gx.InnerGrid(4), gx.Henon(), 12)
    P = t.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon(), row_major=True)
    Q = P.copy().sortColumnMajor()
    assert Q.isSortedColumnMajor() and P == Q
    assert np.array_equal(P.power_iteration(40)[0], Q.power_iteration(40)[0])


def ref_make_symmetric(r, c, v, anti, row_major):
    """CooMatrix.hpp:376-426 restated in numpy: the transposed entries appended, a stable sort,
    every run of equal (row, column) summed from its last entry backwards, halved."""
    R = np.concatenate([r, c])
    C = np.concatenate([c, r])
    V = np.concatenate([v, -v if anti else v])
    o = np.lexsort((C, R)) if row_major else np.lexsort((R, C))
    R, C, V = R[o], C[o], V[o]
    out = []
    i = 0
    while i < len(R):
        j = i
        while j + 1 < len(R) and R[j + 1] == R[i] and C[j + 1] == C[i]:
            j += 1
        acc = V[j]
        for k in range(j - 1, i - 1, -1):
            acc = acc + V[k]
        out.append((R[i], C[i], acc / 2.0))
        i = j + 1
    return (np.array([x[0] for x in out], np.uint32), np.array([x[1] for x in out], np.uint32),
            np.array([x[2] for x in out]))


@pytest.mark.parametrize("anti", [False, True])
def test_make_symmetric_matches_restatement_and_dense_part(anti):
    rng = np.random.default_rng(1700 + int(anti))
    n, m = 400, 3000
    r = rng.integers(0, n, m).astype(np.uint32)
    c = rng.integers(0, n, m).astype(np.uint32)
    v = rng.random(m)
    for repeated in (False, True):
        rr, cc, vv = r.copy(), c.copy(), v.copy()
        if not repeated:  # unique entries, in no order
            keys = np.unique(rr.astype(np.uint64) * n + cc, return_index=True)[1]
            rr, cc, vv = rr[keys], cc[keys], vv[keys]
            p = rng.permutation(len(rr))
            rr, cc, vv = rr[p], cc[p], vv[p]
        for row_major in (False, True):
            print(f"[api] makeSymmetric anti={anti} repeated={repeated} row_major={row_major}", flush=True)
            a = gx.CooMatrix.from_arrays(rr, cc, vv, n, n)
            if row_major:
                a.sortRowMajor()
                o = np.lexsort((cc, rr))
                src = (rr[o], cc[o], vv[o])
            else:
                src = (rr, cc, vv)
            a.makeSymmetric(anti)
            exp = ref_make_symmetric(*src, anti, row_major)
            for x, y in zip(a.download(), exp):
                assert np.array_equal(x, y)
            if not repeated:
                D = np.zeros((n, n))
                D[rr, cc] = vv
                S = (D - D.T) / 2.0 if anti else (D + D.T) / 2.0
                got = np.zeros((n, n))
                x, y, z = a.download()
                got[x, y] = z
                assert np.array_equal(got, S)
