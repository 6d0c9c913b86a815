"""Files of the reference's formats: the tree's JSON file, the legacy binary file of B12a/B12b,
GAIO's box list and the matrix's JSON file. The expected bytes come from a writer in this file
that follows the reference's writers (ImplicitBoxTree.hpp:435-569, CooMatrix.hpp:473-534,
helpFunctions.hpp:100-132) and a reader of the legacy layout (ImplicitBoxTree.hpp:174-257)."""
import os
import struct

import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree

pytestmark = pytest.mark.gpu


def _vec(name, items):
    return '\n  "%s": [' % name + ",".join("\n    " + x for x in items) + "\n  ]"


def ref_tree_json(center, radius, sd, bits, depth, flags):
    s = "{"
    s += '\n  "B12_VERSION_MAJOR": 3,'
    s += '\n  "B12_VERSION_MINOR": 1,'
    s += _vec("center", ["%.17g" % v for v in center]) + ","
    s += _vec("radius", ["%.17g" % v for v in radius]) + ","
    s += _vec("sdScheme", [str(int(v)) for v in sd]) + ","
    s += '\n  "nLeaves": %d,' % len(bits)
    s += _vec("bitPatternVector", [str(int(v)) for v in bits]) + ","
    s += _vec("depthVector", [str(int(v)) for v in depth]) + ","
    s += _vec("flagsVector", [str(int(v)) for v in flags]) + ","
    s += '\n  "nHostBufferBytes": 4294967296'
    s += "\n}"
    return s


def ref_coo_json(nr, nc, rows, cols, vals):
    s = "{"
    s += '\n  "nRows": %d,' % nr
    s += '\n  "nColumns": %d,' % nc
    s += '\n  "nnz": %d,' % len(vals)
    s += _vec("rowIndices", [str(int(v)) for v in rows]) + ","
    s += _vec("columnIndices", [str(int(v)) for v in cols]) + ","
    s += _vec("values", ["%.17g" % v for v in vals]) + ","
    s += "\n}"
    return s


def default_scheme(dim, n):
    return [i % dim for i in range(n)]


@pytest.mark.parametrize("dim,sd,seed", [(1, None, 1), (2, None, 2), (3, None, 3), (4, None, 4), (3, [2, 0, 1, 1], 5)])
def test_tree_json_bytes_and_round_trip(tmp_path, dim, sd, seed):
    print(f"[files] tree json dim={dim} sd={sd}", flush=True)
    rng = np.random.default_rng(700 + seed)
    c, r = rng.uniform(-3, 3, dim), rng.uniform(0.1, 5, dim)
    g = gx.ImplicitBoxTree(c, r, sd)
    g.upload(*random_tree(rng, dim, 500, max_depth=20))
    bits, depth, flags = g.download()
    # the reference keeps 64 scheme entries for its default scheme and 65 for a given one,
    # repeated periodically
    scheme = default_scheme(dim, 64) if sd is None else [sd[i % len(sd)] for i in range(65)]
    path = str(tmp_path / "tree")
    g.save(path)
    assert os.path.exists(path + ".json")
    text = open(path + ".json").read()
    assert text == ref_tree_json(c, r, scheme, bits, depth, flags)
    h = gx.ImplicitBoxTree.load(path, dim)  # the name without ".json" finds the file
    assert same_tree(h.download(), (bits, depth, flags))
    c2, r2, sd2 = h.geometry()
    assert np.array_equal(c2, c) and np.array_equal(r2, r)
    assert list(sd2) == [scheme[i % len(scheme)] for i in range(65)]
    # the same search results on the loaded tree
    pts = rng.uniform(c - r, c + r, (2000, dim))
    assert np.array_equal(g.searchByPoints(pts), h.searchByPoints(pts))


def test_tree_json_merges_equal_boxes_and_reports_missing_files(tmp_path):
    """The loader merges equal consecutive boxes (flags ANDed), as the reference's loader does;
    a missing file is an error."""
    c, r = [0.0, 0.0], [1.0, 1.0]
    bits = [0, 0, 1, 3]
    depth = [1, 1, 2, 2]
    flags = [0x09, 0x08, 0x01, 0x10]
    path = str(tmp_path / "dups.json")
    with open(path, "w") as f:
        f.write(ref_tree_json(c, r, default_scheme(2, 64), bits, depth, flags))
    h = gx.ImplicitBoxTree.load(path, 2)
    b, d, fl = h.download()
    assert list(b) == [0, 1, 3] and list(d) == [1, 2, 2] and list(fl) == [0x08, 0x01, 0x10]
    with pytest.raises(gx.B12xError):
        gx.ImplicitBoxTree.load(str(tmp_path / "nothing_here"), 2)


def test_legacy_binary_file(tmp_path):
    """A B12a/B12b binary file: version words, the dimension as u64, centre, radius, 65 scheme
    bytes, the leaf count, 128-bit patterns (high word first), depths, flags, the host buffer
    size. A leaf deeper than 64 comes back truncated to depth 64."""
    c, r = [0.5, -1.0, 2.0], [0.5, 2.0, 1.0]
    sd = default_scheme(3, 65)
    leaves = [(0, 1, 0x08), (2, 2, 0x01), ((3 << 64) | 5, 66, 0x09)]  # the last: depth 66
    blob = struct.pack("<II", 2, 0) + struct.pack("<Q", 3) + struct.pack("<3d", *c) + struct.pack("<3d", *r)
    blob += bytes(sd) + struct.pack("<Q", len(leaves))
    for b, _, _ in leaves:
        blob += struct.pack("<QQ", b >> 64, b & ((1 << 64) - 1))
    blob += bytes(d for _, d, _ in leaves) + bytes(f for _, _, f in leaves) + struct.pack("<Q", 4 << 30)
    path = str(tmp_path / "old_tree.b12")
    with open(path, "wb") as f:
        f.write(blob)
    h = gx.ImplicitBoxTree.load(path, 3)
    b, d, fl = h.download()
    got = sorted(zip(b.tolist(), d.tolist(), fl.tolist()))
    exp = sorted([(0, 1, 0x08), (2, 2, 0x01), (((3 << 64) | 5) >> 2, 64, 0x09)])
    assert got == exp
    c2, r2, _ = h.geometry()
    assert np.array_equal(c2, c) and np.array_equal(r2, r)


def test_gaio_file(tmp_path):
    rng = np.random.default_rng(710)
    c, r = [0.0, 0.0], [1.5, 0.5]
    g = gx.ImplicitBoxTree(c, r)
    g.upload(*random_tree(rng, 2, 300, max_depth=14))
    path = str(tmp_path / "tree.gaio")
    g.saveForGaio(path)
    raw = open(path, "rb").read()
    dim, sz = struct.unpack_from("<ii", raw, 0)
    assert (dim, sz) == (2, 8)
    off = 8
    cc = struct.unpack_from("<2d", raw, off)
    rr = struct.unpack_from("<2d", raw, off + 16)
    off += 32
    sd = struct.unpack_from("<%di" % (500 * 2 + 1), raw, off)
    off += 4 * (500 * 2 + 1)
    (nb,) = struct.unpack_from("<d", raw, off)
    off += 8
    rows = np.frombuffer(raw, np.float64, count=int(nb) * 4, offset=off).reshape(-1, 4)
    assert off + rows.nbytes == len(raw)
    assert list(cc) == c and list(rr) == r and list(sd) == [i % 64 % 2 for i in range(1001)]
    bits, depth, flags = g.download()
    boxes = g.boxes(-1)
    assert int(nb) == len(bits)
    assert np.array_equal(rows[:, :2], boxes[:, :2])
    assert np.array_equal(rows[:, 2], depth.astype(np.float64)) and np.array_equal(rows[:, 3], flags.astype(np.float64))


def test_matrix_json_bytes_and_round_trip(tmp_path):
    print("[files] matrix json", flush=True)
    g = gx.ImplicitBoxTree([0.0, 0.0], [1.5, 0.5])
    g.setFlags(gx.SD, 0)
    g.subdivision_steps(gx.InnerGrid(4), gx.Henon(), 10)
    m = g.transitionMatrixForLeaves(gx.InnerGrid(4), gx.Henon())
    rows, cols, vals = m.download()
    path = str(tmp_path / "matrix")
    m.save(path)
    text = open(path + ".json").read()
    assert text == ref_coo_json(m.nRows, m.nColumns, rows, cols, vals)
    m2 = gx.CooMatrix.load(path)
    assert (m2.nRows, m2.nColumns, m2.nnz) == (m.nRows, m.nColumns, m.nnz)
    for a, b in zip(m2.download(), (rows, cols, vals)):
        assert np.array_equal(a, b)
