"""Files of the reference's formats: the tree's JSON file, the legacy binary file of B12a/B12b,
GAIO's box list and the matrix's JSON file. The expected tree files come from tests/cpp/
json_stream, which writes them with C++ iostreams as B12's stream helpers do (helpFunctions.hpp:
101-216, ImplicitBoxTree.hpp:436-524: the uint8 arrays sdScheme, depthVector and flagsVector as
one raw byte per element); the matrix file from a writer in this file (CooMatrix.hpp:473-534);
the legacy layout from ImplicitBoxTree.hpp:174-257."""
import os
import struct
import subprocess

import numpy as np
import pytest

import b12_amd as gx
from helpers import random_tree, same_tree

pytestmark = pytest.mark.gpu

JSON_STREAM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "json_stream")
# bytes the reference's reader loses (whitespace, quote) or splits or stops at (comma, ']')
UNREADABLE = {9, 10, 11, 12, 13, 32, 34, 44, 93}


def _vec(name, items):
    return '\n  "%s": [' % name + ",".join("\n    " + x for x in items) + "\n  ]"


def ref_tree_json(center, radius, sd, bits, depth, flags, host_bytes=4 << 30):
    """The tree file as json_stream writes it with iostreams (bytes)."""
    words = [len(center), *["%.17g" % v for v in center], *["%.17g" % v for v in radius], len(sd), *map(int, sd),
             len(bits), *map(int, bits), *map(int, depth), *map(int, flags), int(host_bytes)]
    out = subprocess.run([JSON_STREAM, "write"], input=" ".join(map(str, words)).encode(), capture_output=True,
                         check=True)
    return out.stdout


def readable_tree(rng, dim, n, max_depth):
    """A random tree whose depths and flags the reference's reader reads back."""
    bits, depth, _ = random_tree(rng, dim, n, max_depth=max_depth)
    keep = np.array([int(d) not in UNREADABLE for d in depth])
    flags = rng.choice(np.array([0x00, 0x01, 0x08, 0x10, 0x18, 0x11], np.uint8), len(bits))
    return bits[keep], depth[keep], flags[keep]


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
    g.upload(*readable_tree(rng, dim, 500, max_depth=24))
    bits, depth, flags = g.download()
    # the reference keeps 64 scheme entries for its default scheme and 65 for a given one,
    # repeated periodically
    scheme = default_scheme(dim, 64) if sd is None else [sd[i % len(sd)] for i in range(65)]
    path = str(tmp_path / "tree")
    g.save(path)
    assert os.path.exists(path + ".json")
    raw = open(path + ".json", "rb").read()
    assert raw == ref_tree_json(c, r, scheme, bits, depth, flags)
    h = gx.ImplicitBoxTree.load(path, dim)  # the name without ".json" finds the file
    assert same_tree(h.download(), (bits, depth, flags))
    c2, r2, sd2 = h.geometry()
    assert np.array_equal(c2, c) and np.array_equal(r2, r)
    assert list(sd2) == [scheme[i % len(scheme)] for i in range(65)]
    # the same search results on the loaded tree
    pts = rng.uniform(c - r, c + r, (2000, dim))
    assert np.array_equal(g.searchByPoints(pts), h.searchByPoints(pts))


def test_tree_json_bytes_of_every_depth_and_flag_value(tmp_path):
    """Every depth 1..64 and every flag byte 0..255 is written as its raw byte, as iostreams
    write it; trees holding a byte the reference's reader loses are rejected on loading, with
    the reason, while json_stream's reader (the same extraction) returns fewer elements."""
    print("[files] tree json every byte", flush=True)
    c, r = [0.0], [1.0]
    # pattern 0...01 at every depth d = 1..64: the right half of the lowest box of depth d - 1,
    # so the leaves are disjoint
    bits, depth = [1] * 64, list(range(1, 65))
    flags = [int(v) for v in np.arange(64) * 4 % 256]
    g = gx.ImplicitBoxTree(c, r)
    g.upload(np.array(bits, np.uint64), np.array(depth, np.uint8), np.array(flags, np.uint8))
    b, d, f = g.download()
    path = str(tmp_path / "every.json")
    g.save(path)
    raw = open(path, "rb").read()
    assert raw == ref_tree_json(c, r, default_scheme(1, 64), b, d, f)
    for name, vals in (("depthVector", d), ("flagsVector", f)):
        got = subprocess.run([JSON_STREAM, "read", name], stdin=open(path, "rb"), capture_output=True, check=True)
        back = [int(x) for x in got.stdout.split()]
        assert len(back) < len(vals)  # bytes 9-13, 32, 34, 44 and 93 are among them
    with pytest.raises(gx.B12xError, match="fewer entries than nLeaves"):
        gx.ImplicitBoxTree.load(path, 1)
    # every flag value on its own: written raw, read back unless the reader loses it
    for v in range(256):
        g1 = gx.ImplicitBoxTree(c, r)
        g1.upload(np.array([0, 1], np.uint64), np.array([1, 1], np.uint8), np.array([v, 0], np.uint8))
        p1 = str(tmp_path / ("f%d.json" % v))
        g1.save(p1)
        assert open(p1, "rb").read() == ref_tree_json(c, r, default_scheme(1, 64), [0, 1], [1, 1], [v, 0])
        if v in UNREADABLE:
            with pytest.raises(gx.B12xError):
                gx.ImplicitBoxTree.load(p1, 1)
        else:
            assert list(gx.ImplicitBoxTree.load(p1, 1).download()[2]) == [v, 0]


def test_tree_json_merges_equal_boxes_and_reports_missing_files(tmp_path):
    """The loader merges equal consecutive boxes (flags ANDed), as the reference's loader does;
    a missing file is an error."""
    c, r = [0.0, 0.0], [1.0, 1.0]
    bits = [0, 0, 1, 3]
    depth = [1, 1, 2, 2]
    flags = [0x19, 0x18, 0x01, 0x10]
    path = str(tmp_path / "dups.json")
    with open(path, "wb") as f:
        f.write(ref_tree_json(c, r, default_scheme(2, 64), bits, depth, flags))
    h = gx.ImplicitBoxTree.load(path, 2)
    b, d, fl = h.download()
    assert list(b) == [0, 1, 3] and list(d) == [1, 2, 2] and list(fl) == [0x18, 0x01, 0x10]
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
