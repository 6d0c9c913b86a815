"""ctypes wrapper of the CPU oracle (oracle/liboracle_b12cpu.so) with the same method names as
b12_amd.api.ImplicitBoxTree, so a parity test can run one script against both.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_b12cpu.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libb12ref_maps.so")

_P = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.b12cpu_last_error.restype = ctypes.c_char_p
        _lib.b12cpu_num_points.restype = ctypes.c_uint64
        _lib.b12cpu_coo_nnz.restype = ctypes.c_uint64
        _lib.b12cpu_scheme_points_per_box.restype = ctypes.c_uint64
        _lib.b12cpu_power_iteration.argtypes = [_P, ctypes.c_int, ctypes.c_double, _P, _P, _P]
        _lib.b12cpu_set_options.argtypes = [_P, ctypes.c_int, ctypes.c_uint64]
        _lib.b12cpu_energy_band_stencil.argtypes = [_P, ctypes.c_double, ctypes.c_double, _P]
        _lib.b12cpu_set_threads.argtypes = [ctypes.c_int]
        _lib.b12cpu_subspace_iteration.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _P,
                                                   _P, _P, _P]
        _lib.b12cpu_krylov_create.argtypes = [_P, ctypes.c_int, _P]
        _lib.b12cpu_krylov_destroy.argtypes = [_P]
        _lib.b12cpu_krylov_start.argtypes = [_P, _P, _P]
        _lib.b12cpu_krylov_expand.argtypes = [_P, ctypes.c_int, _P, _P]
        _lib.b12cpu_krylov_rotate.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P]
        _lib.b12cpu_krylov_vector.argtypes = [_P, ctypes.c_int, _P]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _chk(rc):
    if rc != 0:
        raise RuntimeError(lib().b12cpu_last_error().decode())


def _u8(x):
    return ctypes.c_uint8(x & 0xFF)


class OScheme:
    def __init__(self, dim, grid, mp):
        L = lib()
        gp = np.asarray(grid.params, np.uint64)
        mpar = np.asarray(mp.params if mp.params else [0.0], np.float64)
        h = _P()
        _chk(L.b12cpu_scheme_create(dim, grid.kind, _ptr(gp), len(grid.params), mp.kind, _ptr(mpar), len(mp.params),
                                    ctypes.byref(h)))
        self.h, self.dim = h, dim
        self.k = L.b12cpu_scheme_points_per_box(self.h)

    def map_points(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.dim)
        y = np.empty_like(x)
        _chk(lib().b12cpu_map_points(self.h, _ptr(x), _ptr(y), ctypes.c_uint64(x.shape[0])))
        return y

    def grid_points(self, i0, n):
        out = np.empty((n, self.dim), np.float64)
        _chk(lib().b12cpu_grid_points(self.h, ctypes.c_uint64(i0), ctypes.c_uint64(n), _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().b12cpu_scheme_destroy(self.h)


class OIter:
    def __init__(self, scheme, depth, check=0):
        self.scheme, self.depth, self.flags_to_check = scheme, depth & 0xFF, check


class OCoo:
    def __init__(self, h):
        self.h = h
        L = lib()
        self.nnz = L.b12cpu_coo_nnz(h)
        nr, nc = ctypes.c_uint32(), ctypes.c_uint32()
        L.b12cpu_coo_get(h, None, None, None, ctypes.byref(nr), ctypes.byref(nc))
        self.nRows, self.nColumns = nr.value, nc.value

    @classmethod
    def from_arrays(cls, rows, cols, vals, n_rows, n_cols):
        rows = np.ascontiguousarray(rows, np.uint32)
        cols = np.ascontiguousarray(cols, np.uint32)
        vals = np.ascontiguousarray(vals, np.float64)
        h = _P()
        _chk(lib().b12cpu_coo_create(_ptr(rows), _ptr(cols), _ptr(vals), ctypes.c_uint64(len(rows)),
                                     ctypes.c_uint32(n_rows), ctypes.c_uint32(n_cols), ctypes.byref(h)))
        return cls(h)

    def download(self):
        r = np.empty(self.nnz, np.uint32)
        c = np.empty(self.nnz, np.uint32)
        v = np.empty(self.nnz, np.float64)
        nr, nc = ctypes.c_uint32(), ctypes.c_uint32()
        lib().b12cpu_coo_get(self.h, _ptr(r), _ptr(c), _ptr(v), ctypes.byref(nr), ctypes.byref(nc))
        return r, c, v

    def scc(self):
        """Strong components of the matrix's graph: per vertex the largest vertex index of its
        component; also the number of components and of those with more than one vertex."""
        lab = np.empty(self.nRows, np.uint32)
        nc, nb = ctypes.c_uint32(), ctypes.c_uint32()
        _chk(lib().b12cpu_coo_scc(self.h, _ptr(lab), ctypes.byref(nc), ctypes.byref(nb)))
        return lab, nc.value, nb.value

    def power_iteration(self, max_iter=200, tol=-1.0, mu=None):
        m = np.zeros(self.nRows, np.float64) if mu is None else np.ascontiguousarray(mu, np.float64).copy()
        it, res = ctypes.c_int(), ctypes.c_double()
        _chk(lib().b12cpu_power_iteration(self.h, int(max_iter), float(tol), _ptr(m), ctypes.byref(it),
                                          ctypes.byref(res)))
        return m, it.value, res.value

    def subspace_iteration(self, m=4, iters=100, want_basis=False, tol=-1.0, lead=None):
        """(Ritz matrix H (m x m), basis X (m x n) or None, basis updates, residual) of the
        build-defined subspace iteration: at most `iters` updates, stopping once the residual of
        the leading `lead` basis columns (all m by default), ||X_L P - X_L H_LL||_F / ||H_LL||_F,
        is at most tol (tol < 0: exactly `iters` updates)."""
        H = np.zeros((m, m), np.float64)
        X = np.zeros((m, self.nRows), np.float64) if want_basis else None
        it, res = ctypes.c_int(), ctypes.c_double()
        _chk(lib().b12cpu_subspace_iteration(self.h, int(m), int(lead or m), int(iters), float(tol), _ptr(H),
                                             _ptr(X) if want_basis else None, ctypes.byref(it), ctypes.byref(res)))
        return H, X, it.value, res.value

    def krylov(self, ncv):
        """The oracle's Krylov basis (KrylovBasis) for eigs()."""
        return OKrylov(self, ncv)

    def __del__(self):
        if getattr(self, "h", None):
            lib().b12cpu_coo_destroy(self.h)


class OKrylov:
    """The oracle's Krylov basis: the operations of b12_amd.KrylovBasis, restated on the CPU."""

    def __init__(self, matrix, ncv):
        self.n, self.ncv = matrix.nRows, int(ncv)
        h = _P()
        _chk(lib().b12cpu_krylov_create(matrix.h, self.ncv, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().b12cpu_krylov_destroy(self.h)
            self.h = None

    def start(self, v0=None):
        norm = ctypes.c_double()
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, np.float64)
        _chk(lib().b12cpu_krylov_start(self.h, _ptr(v0), ctypes.byref(norm)))
        return norm.value

    def expand(self, p):
        h = np.zeros(p + 1, np.float64)
        beta = ctypes.c_double()
        _chk(lib().b12cpu_krylov_expand(self.h, int(p), _ptr(h), ctypes.byref(beta)))
        return h, beta.value

    def rotate(self, m, k, q):
        q = np.ascontiguousarray(q, np.float64)
        _chk(lib().b12cpu_krylov_rotate(self.h, int(m), int(k), _ptr(q)))

    def vector(self, c):
        out = np.empty(self.n, np.float64)
        _chk(lib().b12cpu_krylov_vector(self.h, int(c), _ptr(out)))
        return out


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops; returns the previous count."""
    return lib().b12cpu_set_threads(int(n))


def compressIndexVector(indices, dim):
    idx = np.ascontiguousarray(indices, np.uint32)
    out = np.empty(dim + 1, np.uint32)
    _chk(lib().b12cpu_compress_index_vector(_ptr(idx), ctypes.c_uint64(len(idx)), ctypes.c_uint32(dim), _ptr(out)))
    return out


class OracleTree:
    """Oracle ImplicitBoxTree with the method names of b12_amd.api.ImplicitBoxTree."""

    def __init__(self, center, radius, sd_scheme=None):
        L = lib()
        c = np.ascontiguousarray(center, np.float64)
        r = np.ascontiguousarray(radius, np.float64)
        self.dim = len(c)
        sd = None if sd_scheme is None else np.ascontiguousarray(sd_scheme, np.uint8)
        h = _P()
        _chk(L.b12cpu_tree_create(self.dim, _ptr(c), _ptr(r), _ptr(sd), 0 if sd is None else len(sd),
                                  ctypes.byref(h)))
        self.h = h
        self._schemes = {}

    def __del__(self):
        if getattr(self, "h", None):
            lib().b12cpu_tree_destroy(self.h)

    def set_options(self, eval_twice=False, chunk_points=0):
        _chk(lib().b12cpu_set_options(self.h, int(eval_twice), int(chunk_points)))

    def _ret(self, name, *args):
        r = ctypes.c_uint32()
        _chk(getattr(lib(), name)(self.h, *args, ctypes.byref(r)))
        return r.value

    def size(self):
        return self._ret("b12cpu_size")

    def count(self, depth=-1):
        return self._ret("b12cpu_count", _u8(depth))

    def depth(self):
        r = ctypes.c_uint8()
        _chk(lib().b12cpu_depth(self.h, ctypes.byref(r)))
        return r.value

    def download(self):
        n = self.size()
        b, d, f = np.empty(n, np.uint64), np.empty(n, np.uint8), np.empty(n, np.uint8)
        _chk(lib().b12cpu_download(self.h, _ptr(b), _ptr(d), _ptr(f)))
        return b, d, f

    def boxes(self, depth=-1):
        """boxes(depth): array of shape (n, 2*dim + 1) = centre, radius, flags per box."""
        n = ctypes.c_uint64()
        _chk(lib().b12cpu_boxes(self.h, _u8(depth), None, ctypes.byref(n)))
        out = np.empty((n.value, 2 * self.dim + 1), np.float64)
        if n.value:
            _chk(lib().b12cpu_boxes(self.h, _u8(depth), _ptr(out), ctypes.byref(n)))
        return out

    def enumerate(self):
        n = ctypes.c_int()
        _chk(lib().b12cpu_enumerate(self.h, None, ctypes.byref(n)))
        out = np.empty(n.value, np.uint32)
        _chk(lib().b12cpu_enumerate(self.h, _ptr(out), ctypes.byref(n)))
        return out

    def upload(self, bits, depth, flags):
        b = np.ascontiguousarray(bits, np.uint64)
        d = np.ascontiguousarray(depth, np.uint8)
        f = np.ascontiguousarray(flags, np.uint8)
        _chk(lib().b12cpu_upload(self.h, ctypes.c_uint32(len(b)), _ptr(b), _ptr(d), _ptr(f)))

    def subdivide(self, flags=8):
        return self._ret("b12cpu_subdivide", _u8(flags))

    def unsubdivide(self, flags=8):
        return self._ret("b12cpu_unsubdivide", _u8(flags))

    def remove(self, flags):
        return self._ret("b12cpu_remove", _u8(flags))

    def setFlags(self, flags, depth=-1):
        return self._ret("b12cpu_set_flags", _u8(flags), _u8(depth))

    def unsetFlags(self, flags, depth=-1):
        return self._ret("b12cpu_unset_flags", _u8(flags), _u8(depth))

    def changeFlags(self, fu, fs, depth=-1):
        return self._ret("b12cpu_change_flags", _u8(fu), _u8(fs), _u8(depth))

    def _stencil(self, op, st, f, f2, depth):
        st = np.ascontiguousarray(np.asarray(st).astype(np.uint8))
        return self._ret("b12cpu_flags_by_stencil", op, _ptr(st), _u8(f), _u8(f2), _u8(depth))

    def setFlagsByStencil(self, st, flags, depth=-1):
        return self._stencil(0, st, flags, 0, depth)

    def unsetFlagsByStencil(self, st, flags, depth=-1):
        return self._stencil(1, st, flags, 0, depth)

    def changeFlagsByStencil(self, st, fu, fs, depth=-1):
        return self._stencil(2, st, fu, fs, depth)

    def energy_band_stencil(self, emin, emax):
        st = np.zeros(self.size(), np.uint8)
        _chk(lib().b12cpu_energy_band_stencil(self.h, float(emin), float(emax), _ptr(st)))
        return st

    def setFlagsByEnergyBand(self, emin, emax, flags, depth=-1):
        return self.setFlagsByStencil(self.energy_band_stencil(emin, emax), flags, depth)

    def scheme(self, grid, mp):
        key = (grid.kind, tuple(grid.params), mp.kind, tuple(mp.params))
        if key not in self._schemes:
            self._schemes[key] = OScheme(self.dim, grid, mp)
        return self._schemes[key]

    def makePointIteratorPairOverLeaves(self, grid, mp):
        return OIter(self.scheme(grid, mp), 255)

    def makePointIteratorPairOverRoot(self, grid, mp):
        return OIter(self.scheme(grid, mp), 0)

    def makePointIteratorPairOverDepth(self, depth, grid, mp):
        return self.makePointIteratorPairByFlags(depth, 0, grid, mp)

    def makePointIteratorPairByFlags(self, depth, check, grid, mp):
        return OIter(self.scheme(grid, mp), depth, check)

    def num_points(self, it):
        return lib().b12cpu_num_points(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check))

    def _by_image(self, it, op, f, f2, depth):
        return self._ret("b12cpu_flags_by_image", it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), op, _u8(f),
                         _u8(f2), _u8(depth))

    def setFlagsByIterators(self, it, flags, depth=-1):
        return self._by_image(it, 0, flags, 0, depth)

    def unsetFlagsByIterators(self, it, flags, depth=-1):
        return self._by_image(it, 1, flags, 0, depth)

    def changeFlagsByIterators(self, it, fu, fs, depth=-1):
        return self._by_image(it, 2, fu, fs, depth)

    def searchByIterators(self, it, depth=-1):
        out = np.empty(self.num_points(it), np.uint32)
        _chk(lib().b12cpu_search_by_image(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), _u8(depth),
                                          _ptr(out)))
        return out

    def searchCompacted(self, it, depth=-1):
        n = ctypes.c_uint64()
        out = np.empty(max(self.size(), 1), np.uint32)
        _chk(lib().b12cpu_search_compacted(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), _u8(depth),
                                           _ptr(out), ctypes.byref(n)))
        return out[: n.value].copy()

    def deleteDepth(self, depth):
        return self._ret("b12cpu_delete_depth", _u8(depth))

    def insertByIterators(self, it, depth, flags_for_inserted=2, flags_for_hit=0):
        res = np.zeros(self.num_points(it), np.uint8)
        _chk(lib().b12cpu_insert_by_image(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), _u8(depth),
                                          _u8(flags_for_inserted), _u8(flags_for_hit), _ptr(res)))
        return res

    def insertByPoints(self, points, depth, flags_for_inserted=2, flags_for_hit=0):
        p = np.ascontiguousarray(points, np.float64).reshape(-1, self.dim)
        res = np.zeros(p.shape[0], np.uint8)
        _chk(lib().b12cpu_insert_by_points(self.h, _ptr(p), ctypes.c_uint64(p.shape[0]), _u8(depth),
                                           _u8(flags_for_inserted), _u8(flags_for_hit), _ptr(res)))
        return res

    def searchByPoints(self, points, depth=-1):
        p = np.ascontiguousarray(points, np.float64).reshape(-1, self.dim)
        out = np.empty(p.shape[0], np.uint32)
        _chk(lib().b12cpu_search_by_points(self.h, _ptr(p), ctypes.c_uint64(p.shape[0]), _u8(depth), _ptr(out)))
        return out

    def point_boxes(self, it):
        n = self.num_points(it)
        b, d = np.empty(n, np.uint64), np.empty(n, np.uint8)
        _chk(lib().b12cpu_point_boxes(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), _ptr(b), _ptr(d)))
        return b, d

    def subdivision_steps(self, grid, mp, steps, sd_flag=8):
        it = None
        counts = np.zeros((steps, 4), np.uint32)
        for s in range(steps):
            counts[s, 0] = self.subdivide(sd_flag)
            it = self.makePointIteratorPairOverLeaves(grid, mp)
            counts[s, 1] = self.setFlagsByIterators(it, 1, -1)
            counts[s, 2] = self.remove(1)
            counts[s, 3] = self.unsetFlags(1, -1)
        return counts

    def transitionMatrix(self, it, depth=-1, use_abs=False, row_major=False, use_invalid=False):
        h = _P()
        _chk(lib().b12cpu_transition_matrix(self.h, it.scheme.h, _u8(it.depth), _u8(it.flags_to_check), _u8(depth),
                                            int(use_abs), int(row_major), int(use_invalid), ctypes.byref(h)))
        return OCoo(h)

    def transitionMatrixForLeaves(self, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverLeaves(grid, mp), -1, use_abs, row_major,
                                     use_invalid)

    def transitionMatrixForDepth(self, depth, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverDepth(depth, grid, mp), depth, use_abs, row_major,
                                     use_invalid)

    def transitionMatrixForRoot(self, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverRoot(grid, mp), 0, use_abs, row_major, use_invalid)


def ref_lib():
    """The reference's own map headers built into oracle/_ref (None when not built)."""
    if not os.path.exists(REF_SO):
        return None
    return ctypes.CDLL(REF_SO)
