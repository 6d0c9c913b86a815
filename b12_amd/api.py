"""Python mirror of B12's ImplicitBoxTree / CooMatrix API over the b12x C ABI.

Names, argument meaning and return values follow the reference's C++ API
(B12c/src/include/ImplicitBoxTree.h, CooMatrix.h, Grids.h, Map_*.h). Everything runs in the
HIP library libb12x.so; there is no CPU fallback — a missing library or GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NONE, HIT, INS, EXPD, SD, EQU, ENTR, EXTR, ALL = 0x00, 0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0xFF
LEAF = 255
INVALID = 0xFFFFFFFF

GRID_INNER, GRID_UNIFORM = 0, 1
MAP_IDENTITY, MAP_HENON, MAP_LORENZ, MAP_HENON_HEILES, MAP_DUFFING = 0, 1, 2, 3, 4

_P = ctypes.c_void_p
_U8 = ctypes.c_uint8
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_I = ctypes.c_int
_D = ctypes.c_double

# name -> argtypes (every entry point returns int status)
SIGNATURES = {
    "b12x_device_name": [ctypes.c_char_p, _I],
    "b12x_tree_create": [_I, _P, _P, _P, _I, _P],
    "b12x_tree_destroy": [_P],
    "b12x_size": [_P, _P],
    "b12x_count": [_P, _U8, _P],
    "b12x_depth": [_P, _P],
    "b12x_boxes": [_P, _U8, _P, _P],
    "b12x_tree_geometry": [_P, _P, _P, _P, _P, _P],
    "b12x_tree_set_domain": [_P, _P, _P, _P, _I, _I],
    "b12x_all_leaves_on_depth": [_P, _U8, _P],
    "b12x_has_same_boxes": [_P, _P, _P],
    "b12x_get_host_buffer_bytes": [_P, _P],
    "b12x_set_host_buffer_bytes": [_P, _U64],
    "b12x_delete_depth": [_P, _U8, _P],
    "b12x_insert_by_image": [_P, _P, _U8, _U8, _U8, _U8, _U8, _P],
    "b12x_insert_by_points": [_P, _P, _U64, _U8, _U8, _U8, _P],
    "b12x_continuation": [_P, _P, _U8, _P, _U64, _P, _I, _P],
    "b12x_tree_save_json": [_P, ctypes.c_char_p],
    "b12x_tree_save_gaio": [_P, ctypes.c_char_p],
    "b12x_tree_load": [ctypes.c_char_p, _I, _P],
    "b12x_coo_save_json": [_P, ctypes.c_char_p],
    "b12x_coo_load_json": [ctypes.c_char_p, _P],
    "b12x_enumerate": [_P, _P, _P],
    "b12x_tree_download": [_P, _P, _P, _P],
    "b12x_tree_upload": [_P, _U32, _P, _P, _P],
    "b12x_subdivide": [_P, _U8, _P],
    "b12x_unsubdivide": [_P, _U8, _P],
    "b12x_remove": [_P, _U8, _P],
    "b12x_set_flags": [_P, _U8, _U8, _P],
    "b12x_unset_flags": [_P, _U8, _U8, _P],
    "b12x_change_flags": [_P, _U8, _U8, _U8, _P],
    "b12x_flags_by_stencil": [_P, _I, _P, _U8, _U8, _U8, _P],
    "b12x_flags_by_energy_band": [_P, _D, _D, _I, _U8, _U8, _U8, _P],
    "b12x_flags_by_image": [_P, _P, _U8, _U8, _I, _U8, _U8, _U8, _P],
    "b12x_num_points": [_P, _P, _U8, _U8, _P],
    "b12x_search_by_image": [_P, _P, _U8, _U8, _U8, _P],
    "b12x_search_compacted": [_P, _P, _U8, _U8, _U8, _P, _P],
    "b12x_search_by_points": [_P, _P, _U64, _U8, _P],
    "b12x_subdivision_steps": [_P, _P, _U8, _I, _P],
    "b12x_tree_image_timing": [_P, _I, _P, _P, _P],
    "b12x_tree_image_launches": [_P, _P, _P, _U64, _P],
    "b12x_set_tm_pass_bytes": [_P, _U64],
    "b12x_get_tm_pass_bytes": [_P, _P],
    "b12x_tm_last_passes": [_P],
    "b12x_bench_copy": [ctypes.c_uint64, _I, _P],
    "b12x_bench_fp64_clock": [_P, _P],
    "b12x_scheme_create": [_I, _I, _P, _I, _I, _P, _I, _P],
    "b12x_scheme_destroy": [_P],
    "b12x_scheme_points_per_box": [_P, _P],
    "b12x_map_points": [_P, _P, _P, _U64],
    "b12x_transition_matrix": [_P, _P, _U8, _U8, _U8, _I, _I, _I, _P],
    "b12x_coo_destroy": [_P],
    "b12x_coo_info": [_P, _P, _P, _P, _P],
    "b12x_coo_download": [_P, _P, _P, _P],
    "b12x_coo_upload": [_P, _P, _P, _U64, _U32, _U32, _P],
    "b12x_compress_index_vector": [_P, _U64, _U32, _P],
    "b12x_coo_is_sorted": [_P, _I, _P],
    "b12x_coo_sort": [_P, _I],
    "b12x_coo_copy": [_P, _P],
    "b12x_coo_make_symmetric": [_P, _I],
    "b12x_power_iteration": [_P, _I, _D, _I, _P, _P, _P],
    "b12x_power_iteration_kernel_ms": [_P],
    "b12x_power_iteration_form": [_P],
    "b12x_power_iteration_noalloc": [_P, _I, _D, _I, _P, _P, _P],
    "b12x_subspace_iteration": [_P, _I, _I, _I, _D, _P, _P, _P, _P],
    "b12x_krylov_create": [_P, _I, _P],
    "b12x_krylov_destroy": [_P],
    "b12x_krylov_start": [_P, _P, _P],
    "b12x_krylov_expand": [_P, _I, _P, _P],
    "b12x_krylov_rotate": [_P, _I, _I, _P],
    "b12x_krylov_vector": [_P, _I, _P],
    "b12x_coo_scc": [_P, _P, _P, _P],
    "b12x_scc_last_stats": [_P, _P, _P],
    "b12x_chain_recurrent_steps": [_P, _P, _U8, _I, _P],
    "b12x_forest_create": [_I, _I, _P, _I, _I, _I, _I, _I, _D, _P],
    "b12x_forest_destroy": [_P],
    "b12x_forest_add": [_P, _P, _P, _P, _I, _U8, _P],
    "b12x_forest_step": [_P, _P, _P],
    "b12x_forest_run": [_P],
    "b12x_forest_job_info": [_P, _U32, _P, _P, _P, _P, _P],
    "b12x_forest_job_cover": [_P, _U32, _P, _P, _P],
    "b12x_forest_job_matrix": [_P, _U32, _P, _P, _P],
    "b12x_forest_job_measure": [_P, _U32, _P],
    "b12x_forest_job_release": [_P, _U32],
    "b12x_forest_footprint": [_P, _P],
    "b12x_forest_stats": [_P, _P, ctypes.c_int],
    "b12x_forest_set_pool_iterations": [_P, ctypes.c_int],
    "b12x_serve_run": [_P, _I, _U32, _P, _P, _P, _P, _I, _P, _I, _P, _D, _D, _P, _P, _P, _P, _P],
}

_lib = None


def library_path(flavour="parity"):
    name = "libb12x.so" if flavour == "parity" else "libb12x_fast.so"
    return os.path.join(HERE, name)


def load(flavour="parity"):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None and flavour == "parity":
        return _lib
    path = library_path(flavour)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m b12_amd.build` (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    lib.b12x_last_error.restype = ctypes.c_char_p
    lib.b12x_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _I
    if flavour == "parity":
        _lib = lib
    return lib


class B12xError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


class _Api:
    def __init__(self, flavour="parity"):
        self.lib = load(flavour)

    def __call__(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            raise B12xError(f"{name}: {self.lib.b12x_last_error().decode()} (status {rc})")


def bench_copy(nbytes, reps=20, flavour="parity"):
    """Average milliseconds of a plain device copy of nbytes (b12x_bench_copy)."""
    ms = _D()
    _Api(flavour)("b12x_bench_copy", int(nbytes), int(reps), ctypes.byref(ms))
    return ms.value


def bench_fp64_clock(flavour="parity"):
    """(shader clock in GHz under a full FP64 load, fraction of the FP64 issue slots used at that
    clock) from b12x_bench_fp64_clock."""
    ghz, frac = _D(), _D()
    _Api(flavour)("b12x_bench_fp64_clock", ctypes.byref(ghz), ctypes.byref(frac))
    return ghz.value, frac.value


def tm_last_passes(flavour="parity"):
    """Passes of this thread's last transitionMatrix."""
    v = _U32()
    _Api(flavour)("b12x_tm_last_passes", ctypes.byref(v))
    return v.value


def power_iteration_kernel_ms(flavour="parity"):
    """Duration (ms) of this thread's last persistent power-iteration launch (0: batched form)."""
    v = _D()
    _Api(flavour)("b12x_power_iteration_kernel_ms", ctypes.byref(v))
    return v.value


PI_FORMS = {-1: "batched", 0: "straight", 1: "sweep"}


def power_iteration_form(flavour="parity"):
    """Form of this thread's last power iteration: "straight" (the persistent form, one group of
    64 columns per wave), "sweep" (the persistent form, waves sweeping several groups) or
    "batched" (the batched form alone)."""
    v = ctypes.c_int()
    _Api(flavour)("b12x_power_iteration_form", ctypes.byref(v))
    return PI_FORMS[v.value]


def scc_last_stats(flavour="parity"):
    """(rounds, passes, host checks) of this thread's last strong-components run."""
    v = [_U32(), _U32(), _U32()]
    _Api(flavour)("b12x_scc_last_stats", *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def device_name(flavour="parity"):
    buf = ctypes.create_string_buffer(256)
    _Api(flavour)("b12x_device_name", buf, 256)
    return buf.value.decode()


# ---- functors (Grids.h, Map_*.h): kind + parameters, compiled into the library --------------
class InnerGrid:
    """InnerGrid<DIM>(nPointsPerDim) — Grids.hpp:99-134"""

    def __init__(self, n=1):
        self.kind, self.params = GRID_INNER, [int(n)]
        self.n = int(n)

    def points_per_box(self, dim):
        return self.n ** dim


class UniformGrid:
    """UniformGrid<DIM>(nPoints, seed), host (std::mt19937) semantics — Grids.hpp:136-176"""

    def __init__(self, k=1, seed=0):
        self.kind, self.params = GRID_UNIFORM, [int(k), int(seed)]

    def points_per_box(self, dim):
        return self.params[0]


class _Map:
    def __init__(self, kind, dim, params):
        self.kind, self.dim, self.params = kind, dim, [float(p) for p in params]


def Identity(dim):
    return _Map(MAP_IDENTITY, dim, [])


def Henon(a=1.4, b=0.3, steps=1):
    return _Map(MAP_HENON, 2, [a, b, steps])


def Lorenz(sigma=10.0, rho=28.0, beta=8.0 / 3.0, h=0.01, steps=20):
    return _Map(MAP_LORENZ, 3, [sigma, rho, beta, h, steps])


def HenonHeiles(order=2, h=0.02, steps=50):
    return _Map(MAP_HENON_HEILES, 4, [order, h, steps])


def Duffing(delta=0.25, alpha=-1.0, beta=1.0, gamma=0.3, omega=1.0, t0=0.0, t_end=2 * np.pi, h=2 * np.pi / 100):
    return _Map(MAP_DUFFING, 2, [delta, alpha, beta, gamma, omega, t0, t_end, h])


_SCHEMES = {}  # (dim, grid kind, grid params, map kind, map params, flavour) -> Scheme


class Scheme:
    """A (Grid, Map) pair on the device."""

    def __init__(self, dim, grid, mp, flavour="parity"):
        self.api = _Api(flavour)
        gp = np.asarray(grid.params, dtype=np.uint64)
        mpar = np.asarray(mp.params if mp.params else [0.0], dtype=np.float64)
        h = _P()
        self.api("b12x_scheme_create", dim, grid.kind, _ptr(gp), len(grid.params), mp.kind, _ptr(mpar),
                 len(mp.params), ctypes.byref(h))
        self.h = h
        self.dim = dim
        k = _U64()
        self.api("b12x_scheme_points_per_box", self.h, ctypes.byref(k))
        self.k = k.value

    def map_points(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, self.dim)
        y = np.empty_like(x)
        self.api("b12x_map_points", self.h, _ptr(x), _ptr(y), x.shape[0])
        return y

    def __del__(self):
        if getattr(self, "h", None):
            self.api.lib.b12x_scheme_destroy(self.h)
            self.h = None


class PointIterator:
    """The lazy point range of makePointIteratorPair* (ImplicitBoxTree.h:246-430)."""

    def __init__(self, scheme, depth, flags_to_check=NONE):
        self.scheme, self.depth, self.flags_to_check = scheme, depth & 0xFF, flags_to_check


class CooMatrix:
    """CooMatrix<A, double> (CooMatrix.h:56-121) with its arrays resident on the GPU."""

    def __init__(self, handle, api):
        self.h, self.api = handle, api
        nnz, nr, nc, rm = _U64(), _U32(), _U32(), _I()
        self.api("b12x_coo_info", self.h, ctypes.byref(nnz), ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(rm))
        self.nnz, self.nRows, self.nColumns, self.row_major = nnz.value, nr.value, nc.value, bool(rm.value)

    @classmethod
    def from_arrays(cls, rows, cols, vals, n_rows, n_cols, flavour="parity"):
        api = _Api(flavour)
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        h = _P()
        api("b12x_coo_upload", _ptr(rows), _ptr(cols), _ptr(vals), len(rows), n_rows, n_cols, ctypes.byref(h))
        return cls(h, api)

    @classmethod
    def load(cls, file_name, flavour="parity"):
        """CooMatrix(fileName) (CooMatrix.h:75-76): a matrix saved by save()."""
        api = _Api(flavour)
        h = _P()
        api("b12x_coo_load_json", os.fsencode(file_name), ctypes.byref(h))
        return cls(h, api)

    def save(self, file_name):
        """save(fileName) (CooMatrix.h:110-112): B12's JSON file (".json" appended if missing)."""
        self.api("b12x_coo_save_json", self.h, os.fsencode(file_name))

    def download(self):
        r = np.empty(self.nnz, np.uint32)
        c = np.empty(self.nnz, np.uint32)
        v = np.empty(self.nnz, np.float64)
        self.api("b12x_coo_download", self.h, _ptr(r), _ptr(c), _ptr(v))
        return r, c, v

    def power_iteration(self, max_iter=200, tol=-1.0, mu=None):
        if mu is not None and np.shape(mu) != (self.nRows,):
            raise ValueError(f"mu must hold nRows = {self.nRows} values, got shape {np.shape(mu)}")
        mu_arr = np.zeros(self.nRows, np.float64) if mu is None else np.ascontiguousarray(mu, np.float64).copy()
        it, res = _I(), _D()
        self.api("b12x_power_iteration", self.h, int(max_iter), float(tol), 0 if mu is None else 1, _ptr(mu_arr),
                 ctypes.byref(it), ctypes.byref(res))
        return mu_arr, it.value, res.value

    def power_iteration_noalloc(self, max_iter=200, tol=-1.0, mu=None):
        """power_iteration through b12x_power_iteration_noalloc: no device allocation (a fixed device
        scratch array; column-major matrices of at most 131072 columns)."""
        if mu is not None and np.shape(mu) != (self.nRows,):
            raise ValueError(f"mu must hold nRows = {self.nRows} values, got shape {np.shape(mu)}")
        mu_arr = np.zeros(self.nRows, np.float64) if mu is None else np.ascontiguousarray(mu, np.float64).copy()
        it, res = _I(), _D()
        self.api("b12x_power_iteration_noalloc", self.h, int(max_iter), float(tol), 0 if mu is None else 1,
                 _ptr(mu_arr), ctypes.byref(it), ctypes.byref(res))
        return mu_arr, it.value, res.value

    def subspace_iteration(self, m=4, iters=100, want_basis=False, tol=-1.0, lead=None):
        """(Ritz matrix H (m x m), basis X (m x n) or None, basis updates, residual) of
        b12x_subspace_iteration: at most `iters` updates, stopping once the residual of the
        leading `lead` basis columns (all m by default), ||X_L P - X_L H_LL||_F / ||H_LL||_F, is at
        most tol (tol < 0: exactly `iters` updates); the eigenvalues of H approximate the m leading
        eigenvalues of the matrix."""
        H = np.zeros((m, m), np.float64)
        X = np.zeros((m, self.nRows), np.float64) if want_basis else None
        it, res = _I(), _D()
        self.api("b12x_subspace_iteration", self.h, int(m), int(lead or m), int(iters), float(tol), _ptr(H),
                 _ptr(X) if want_basis else None, ctypes.byref(it), ctypes.byref(res))
        return H, X, it.value, res.value

    def krylov(self, ncv):
        """A Krylov basis of ncv + 1 vectors on the GPU for eigs() (b12x_krylov_*)."""
        return KrylovBasis(self, ncv)

    # ---- ordering and comparison (CooMatrix.h:61-91, CooMatrix.hpp:304-447)
    def isSortedRowMajor(self):
        r = _I()
        self.api("b12x_coo_is_sorted", self.h, 1, ctypes.byref(r))
        return bool(r.value)

    def isSortedColumnMajor(self):
        r = _I()
        self.api("b12x_coo_is_sorted", self.h, 0, ctypes.byref(r))
        return bool(r.value)

    def sortRowMajor(self):
        self.api("b12x_coo_sort", self.h, 1)
        return self

    def sortColumnMajor(self):
        self.api("b12x_coo_sort", self.h, 0)
        return self

    def copy(self):
        h = _P()
        self.api("b12x_coo_copy", self.h, ctypes.byref(h))
        return CooMatrix(h, self.api)

    def makeSymmetric(self, makeAntisymmetric=False):
        """(P + P^T)/2, or (P - P^T)/2 (CooMatrix.hpp:376-426), in place."""
        self.api("b12x_coo_make_symmetric", self.h, int(bool(makeAntisymmetric)))
        nnz = _U64()
        self.api("b12x_coo_info", self.h, ctypes.byref(nnz), None, None, None)
        self.nnz = nnz.value

    def _same_entries(self, other):
        # the reference compares this matrix's entries with the first entries of the other
        if other.nnz < self.nnz:
            return False
        a, b = self.download(), other.download()
        return all(np.array_equal(x, y[: self.nnz]) for x, y in zip(a, b))

    def __eq__(self, other):
        """operator== (CooMatrix.hpp:304-352): the same sizes and entries, independent of the
        orderings (an unsorted or differently sorted matrix is compared through a sorted copy)."""
        if not isinstance(other, CooMatrix):
            return NotImplemented
        if (self.nRows, self.nColumns) != (other.nRows, other.nColumns):
            return False
        frm = self.isSortedRowMajor()
        fcm = not frm and self.isSortedColumnMajor()
        srm = other.isSortedRowMajor()
        scm = not srm and other.isSortedColumnMajor()
        if frm and not srm:
            return self == other.copy().sortRowMajor()
        if fcm and not scm:
            return self == other.copy().sortColumnMajor()
        if srm and not frm:
            return self.copy().sortRowMajor() == other
        if scm and not fcm:
            return self.copy().sortColumnMajor() == other
        if self._same_entries(other):
            return True
        if not ((frm and srm) or (fcm and scm)):
            return self.copy().sortRowMajor() == other.copy().sortRowMajor()
        return False

    __hash__ = None

    def scc(self):
        """Strong components of the matrix's graph (b12x_coo_scc): per vertex the largest vertex
        index of its component, the number of components and of those with more than one vertex."""
        lab = np.empty(self.nRows, np.uint32)
        nc, nb = ctypes.c_uint32(), ctypes.c_uint32()
        self.api("b12x_coo_scc", self.h, _ptr(lab), ctypes.byref(nc), ctypes.byref(nb))
        return lab, nc.value, nb.value

    def __del__(self):
        if getattr(self, "h", None):
            self.api.lib.b12x_coo_destroy(self.h)
            self.h = None


def compressIndexVector(indices, dim, flavour="parity"):
    """compressIndexVector<A>(indices, dim) — CooMatrix.hpp:14-56"""
    api = _Api(flavour)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.empty(dim + 1, np.uint32)
    api("b12x_compress_index_vector", _ptr(idx), len(idx), dim, _ptr(out))
    return out


def _depth(d):
    return d & 0xFF


class ImplicitBoxTree:
    """ImplicitBoxTree<MI355X, DIM, double, 64> — ImplicitBoxTree.h:32-640."""

    def __init__(self, center, radius, sd_scheme=None, flavour="parity"):
        self.api = _Api(flavour)
        self.flavour = flavour
        c = np.ascontiguousarray(center, dtype=np.float64)
        r = np.ascontiguousarray(radius, dtype=np.float64)
        self.dim = len(c)
        sd = None if sd_scheme is None else np.ascontiguousarray(sd_scheme, dtype=np.uint8)
        h = _P()
        self.api("b12x_tree_create", self.dim, _ptr(c), _ptr(r), _ptr(sd), 0 if sd is None else len(sd),
                 ctypes.byref(h))
        self.h = h

    @classmethod
    def load(cls, file_name, dim, flavour="parity"):
        """ImplicitBoxTree(fileName) (ImplicitBoxTree.h:58-59): a tree saved by save(), or a
        legacy binary file of B12a/B12b (a name without ".json")."""
        obj = cls.__new__(cls)
        obj.api = _Api(flavour)
        obj.flavour = flavour
        obj.dim = int(dim)
        h = _P()
        obj.api("b12x_tree_load", os.fsencode(file_name), obj.dim, ctypes.byref(h))
        obj.h = h
        return obj

    def save(self, file_name):
        """save(fileName) (ImplicitBoxTree.h:70-72): B12's JSON file (".json" appended if missing)."""
        self.api("b12x_tree_save_json", self.h, os.fsencode(file_name))

    def saveForGaio(self, file_name):
        """saveForGaio(fileName) (ImplicitBoxTree.h:74-76): GAIO's binary box list."""
        self.api("b12x_tree_save_gaio", self.h, os.fsencode(file_name))

    def geometry(self):
        """(centre, radius, sdScheme) of the tree."""
        d = _I()
        c, r = np.empty(4, np.float64), np.empty(4, np.float64)
        sd, n = np.empty(65, np.uint8), _I()
        self.api("b12x_tree_geometry", self.h, ctypes.byref(d), _ptr(c), _ptr(r), _ptr(sd), ctypes.byref(n))
        return c[:d.value].copy(), r[:d.value].copy(), sd[:n.value].copy()

    def __del__(self):
        if getattr(self, "h", None):
            self.api.lib.b12x_tree_destroy(self.h)
            self.h = None

    def _ret(self, name, *args):
        r = _U32()
        self.api(name, self.h, *args, ctypes.byref(r))
        return r.value

    # -- queries
    def size(self):
        return self._ret("b12x_size")

    def count(self, depth=-1):
        return self._ret("b12x_count", _depth(depth))

    def depth(self):
        r = _U8()
        self.api("b12x_depth", self.h, ctypes.byref(r))
        return r.value

    def boxes(self, depth=-1):
        """boxes(depth) (ImplicitBoxTree.h:86-91): array of shape (n, 2*dim + 1) holding the
        centre, radius and flags of every box at depth, in leaf order."""
        n = ctypes.c_uint64()
        self.api("b12x_boxes", self.h, _depth(depth), None, ctypes.byref(n))
        out = np.empty((n.value, 2 * self.dim + 1), np.float64)
        if n.value:
            self.api("b12x_boxes", self.h, _depth(depth), out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n))
        return out

    def enumerate(self):
        """enumerate() (ImplicitBoxTree.h:132-133): {count(0), ..., count(depth())}."""
        n = ctypes.c_int()
        self.api("b12x_enumerate", self.h, None, ctypes.byref(n))
        out = np.empty(n.value, np.uint32)
        self.api("b12x_enumerate", self.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n))
        return out

    def download(self):
        n = self.size()
        b = np.empty(n, np.uint64)
        d = np.empty(n, np.uint8)
        f = np.empty(n, np.uint8)
        self.api("b12x_tree_download", self.h, _ptr(b), _ptr(d), _ptr(f))
        return b, d, f

    def upload(self, bits, depth, flags):
        b = np.ascontiguousarray(bits, np.uint64)
        d = np.ascontiguousarray(depth, np.uint8)
        f = np.ascontiguousarray(flags, np.uint8)
        self.api("b12x_tree_upload", self.h, len(b), _ptr(b), _ptr(d), _ptr(f))

    def image_timing(self, enable):
        """(ms, launches, points) of image-kernel launches since timing was last enabled."""
        ms, nl, npt = _D(), _U64(), _U64()
        self.api("b12x_tree_image_timing", self.h, int(enable), ctypes.byref(ms), ctypes.byref(nl),
                 ctypes.byref(npt))
        return ms.value, nl.value, npt.value

    def image_launches(self):
        """[(ms, points)] of each image launch timed since timing was last enabled, in order."""
        n = _U64()
        self.api("b12x_tree_image_launches", self.h, None, None, 0, ctypes.byref(n))
        ms = np.zeros(n.value, np.float64)
        pts = np.zeros(n.value, np.uint64)
        self.api("b12x_tree_image_launches", self.h, _ptr(ms), _ptr(pts), n.value, ctypes.byref(n))
        return list(zip(ms.tolist(), pts.tolist()))

    # -- structural operations
    def subdivide(self, flags=SD):
        return self._ret("b12x_subdivide", flags)

    def unsubdivide(self, flags=SD):
        return self._ret("b12x_unsubdivide", flags)

    def remove(self, flags):
        return self._ret("b12x_remove", flags)

    # -- flags
    def setFlags(self, flags, depth=-1):
        return self._ret("b12x_set_flags", flags, _depth(depth))

    def unsetFlags(self, flags, depth=-1):
        return self._ret("b12x_unset_flags", flags, _depth(depth))

    def changeFlags(self, flags_to_unset, flags_to_set, depth=-1):
        return self._ret("b12x_change_flags", flags_to_unset, flags_to_set, _depth(depth))

    def _stencil(self, op, stencil, f, f2, depth):
        st = np.ascontiguousarray(np.asarray(stencil).astype(np.uint8))
        return self._ret("b12x_flags_by_stencil", op, _ptr(st), f, f2, _depth(depth))

    def setFlagsByStencil(self, stencil, flags, depth=-1):
        return self._stencil(0, stencil, flags, 0, depth)

    def setFlagsByEnergyBand(self, emin, emax, flags, depth=-1):
        """Hénon-Heiles energy-band predicate (4-dimensional trees): flags on the leaves whose
        box meets [emin, emax] (b12x_flags_by_energy_band)."""
        return self._ret("b12x_flags_by_energy_band", float(emin), float(emax), 0, flags, 0, _depth(depth))

    def unsetFlagsByStencil(self, stencil, flags, depth=-1):
        return self._stencil(1, stencil, flags, 0, depth)

    def changeFlagsByStencil(self, stencil, flags_to_unset, flags_to_set, depth=-1):
        return self._stencil(2, stencil, flags_to_unset, flags_to_set, depth)

    # ---- by points and by search vector (ImplicitBoxTree.h:97-108, 472-480, 532-539): the boxes of
    # depth `depth` holding the points, or named by a sorted search vector (searchCompacted's
    # output; invalid entries are skipped), select the boxes as a stencil would
    def _search_stencil(self, search, depth):
        st = np.zeros(self.count(depth), np.uint8)
        search = np.asarray(search, np.uint32)
        ok = search[search != INVALID]
        st[ok[ok < len(st)]] = 1
        return st

    def setFlagsByPoints(self, points, flags, depth=-1):
        return self.setFlagsByStencil(self._search_stencil(self.searchByPoints(points, depth), depth), flags, depth)

    def unsetFlagsByPoints(self, points, flags, depth=-1):
        return self.unsetFlagsByStencil(self._search_stencil(self.searchByPoints(points, depth), depth), flags, depth)

    def changeFlagsByPoints(self, points, flags_to_unset, flags_to_set, depth=-1):
        return self.changeFlagsByStencil(self._search_stencil(self.searchByPoints(points, depth), depth),
                                         flags_to_unset, flags_to_set, depth)

    def setFlagsBySearch(self, search, flags, depth=-1):
        return self.setFlagsByStencil(self._search_stencil(search, depth), flags, depth)

    def unsetFlagsBySearch(self, search, flags, depth=-1):
        return self.unsetFlagsByStencil(self._search_stencil(search, depth), flags, depth)

    def changeFlagsBySearch(self, search, flags_to_unset, flags_to_set, depth=-1):
        return self.changeFlagsByStencil(self._search_stencil(search, depth), flags_to_unset, flags_to_set, depth)

    # ---- domain and bookkeeping (ImplicitBoxTree.h:82-85, 141-143, 244, 550-569)
    def getCenter(self):
        return self.geometry()[0]

    def getRadius(self):
        return self.geometry()[1]

    def getSdScheme(self):
        return self.geometry()[2]

    def setCenter(self, center):
        """setCenter: cut to dim entries, or extended with 0.5 (ImplicitBoxTree.hpp:3470-3478)."""
        c = np.full(self.dim, 0.5)
        v = np.asarray(center, np.float64).ravel()[: self.dim]
        c[: len(v)] = v
        self.api("b12x_tree_set_domain", self.h, _ptr(c), None, None, 0, 0)
        return self.getCenter()

    def setRadius(self, radius):
        """setRadius: cut or extended with 0.5; negative entries leave the radius unchanged with a
        warning (ImplicitBoxTree.hpp:3487-3505)."""
        r = np.full(self.dim, 0.5)
        v = np.asarray(radius, np.float64).ravel()[: self.dim]
        r[: len(v)] = v
        if np.any(v < 0):
            print("Warning: Input has negative values. Radius vector remains unchanged.")
            return self.getRadius()
        self.api("b12x_tree_set_domain", self.h, None, _ptr(r), None, 0, 0)
        return self.getRadius()

    def setSdScheme(self, sd_scheme):
        """setSdScheme: repeated periodically, the default scheme when empty; values >= dim leave it
        unchanged with a warning (ImplicitBoxTree.hpp:3513-3552)."""
        sd = np.ascontiguousarray(sd_scheme, np.int64).ravel()
        if np.any(sd < 0) or np.any(sd >= self.dim):
            print("Warning: Input has values >= DIM. SdScheme remains unchanged.")
            return self.getSdScheme()
        sd8 = sd.astype(np.uint8)
        self.api("b12x_tree_set_domain", self.h, None, None, _ptr(sd8) if len(sd8) else None, len(sd8), 1)
        return self.getSdScheme()

    def isSdSchemeDefault(self):
        sd = self.getSdScheme()
        return bool(np.all(sd == np.arange(len(sd)) % self.dim))

    def getSdCount(self):
        """sdCount_ (adaptSdCount, ImplicitBoxTree.hpp:574-587): row i + 1 counts, per coordinate,
        the bisections of the first i + 1 scheme entries; (64 + 2) x dim entries."""
        sd = self.getSdScheme()
        c = np.zeros(((64 + 2), self.dim), np.uint8)
        for i, j in enumerate(sd):
            c[i + 1] = c[i]
            c[i + 1, j] += 1
        return c.ravel()

    def areAllLeavesOnDepth(self, depth):
        r = _I()
        self.api("b12x_all_leaves_on_depth", self.h, _depth(depth), ctypes.byref(r))
        return bool(r.value)

    def hasSameBoxes(self, other):
        r = _I()
        self.api("b12x_has_same_boxes", self.h, other.h, ctypes.byref(r))
        return bool(r.value)

    def getNHostBufferBytes(self):
        r = _U64()
        self.api("b12x_get_host_buffer_bytes", self.h, ctypes.byref(r))
        return r.value

    def setNHostBufferBytes(self, n):
        self.api("b12x_set_host_buffer_bytes", self.h, int(n))
        return self.getNHostBufferBytes()

    def set_tm_pass_bytes(self, n):
        """Per-point scratch bound of a transitionMatrix pass (0: a third of the free HBM)."""
        self.api("b12x_set_tm_pass_bytes", self.h, int(n))

    def get_tm_pass_bytes(self):
        r = _U64()
        self.api("b12x_get_tm_pass_bytes", self.h, ctypes.byref(r))
        return r.value

    def getDepthVector(self):
        return self.download()[1]

    def getFlagsVector(self):
        return self.download()[2]

    def copy(self):
        """A new tree with this tree's domain, scheme, leaves and host buffer size (the
        reference's copy constructor, ImplicitBoxTree.hpp:47-84)."""
        c, r, sd = self.geometry()
        t = ImplicitBoxTree(c, r, None if self.isSdSchemeDefault() and len(sd) == 64 else sd, flavour=self.flavour)
        t.upload(*self.download())
        t.setNHostBufferBytes(self.getNHostBufferBytes())
        return t

    # -- point iterators
    def scheme(self, grid, mp):
        # schemes are immutable device tables: one per (dimension, grid, map, library) for the
        # process, so a tree made per job does not rebuild them
        key = (self.dim, grid.kind, tuple(grid.params), mp.kind, tuple(mp.params), self.flavour)
        sch = _SCHEMES.get(key)
        if sch is None:
            sch = _SCHEMES[key] = Scheme(self.dim, grid, mp, self.flavour)
        return sch

    def makePointIteratorPairOverLeaves(self, grid, mp):
        return PointIterator(self.scheme(grid, mp), LEAF)

    def makePointIteratorPairOverRoot(self, grid, mp):
        return PointIterator(self.scheme(grid, mp), 0)

    def makePointIteratorPairOverDepth(self, depth, grid, mp):
        return self.makePointIteratorPairByFlags(depth, 0, grid, mp)

    def makePointIteratorPairByFlags(self, depth, flags_to_check, grid, mp):
        return PointIterator(self.scheme(grid, mp), depth, flags_to_check)

    def num_points(self, it):
        r = _U64()
        self.api("b12x_num_points", self.h, it.scheme.h, it.depth, it.flags_to_check, ctypes.byref(r))
        return r.value

    def _by_image(self, it, op, f, f2, depth):
        return self._ret("b12x_flags_by_image", it.scheme.h, it.depth, it.flags_to_check, op, f, f2, _depth(depth))

    def setFlagsByIterators(self, it, flags, depth=-1):
        return self._by_image(it, 0, flags, 0, depth)

    def unsetFlagsByIterators(self, it, flags, depth=-1):
        return self._by_image(it, 1, flags, 0, depth)

    def changeFlagsByIterators(self, it, flags_to_unset, flags_to_set, depth=-1):
        return self._by_image(it, 2, flags_to_unset, flags_to_set, depth)

    def searchByIterators(self, it, depth=-1):
        out = np.empty(self.num_points(it), np.uint32)
        self.api("b12x_search_by_image", self.h, it.scheme.h, it.depth, it.flags_to_check, _depth(depth), _ptr(out))
        return out

    def searchCompacted(self, it, depth=-1):
        n = _U64()
        out = np.empty(max(self.size(), 1), np.uint32)
        self.api("b12x_search_compacted", self.h, it.scheme.h, it.depth, it.flags_to_check, _depth(depth), _ptr(out),
                 ctypes.byref(n))
        return out[: n.value].copy()

    def deleteDepth(self, depth):
        """deleteDepth(depth) (ImplicitBoxTree.h:124-126)."""
        return self._ret("b12x_delete_depth", _depth(depth))

    def insertByIterators(self, it, depth, flags_for_inserted=INS, flags_for_hit=NONE):
        """insertByIterators(points, depth, flagsForInserted, flagsForHit) (ImplicitBoxTree.h:170-187):
        isToInsert per point."""
        res = np.zeros(self.num_points(it), np.uint8)
        self.api("b12x_insert_by_image", self.h, it.scheme.h, _depth(it.depth), it.flags_to_check, _depth(depth),
                 flags_for_inserted, flags_for_hit, _ptr(res))
        return res

    def insertByPoints(self, points, depth, flags_for_inserted=INS, flags_for_hit=NONE):
        """insertByPoints(points, depth, flagsForInserted, flagsForHit) (ImplicitBoxTree.h:189-202)."""
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.dim)
        res = np.zeros(p.shape[0], np.uint8)
        self.api("b12x_insert_by_points", self.h, _ptr(p), p.shape[0], _depth(depth), flags_for_inserted,
                 flags_for_hit, _ptr(res))
        return res

    def continuation_run(self, grid, mp, depth, points):
        """continuation.m at one depth on the device (b12x_continuation): the list of count(depth)
        after the first insertion and after every expansion step."""
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.dim)
        cap = 4096
        while True:
            counts = np.zeros(cap, np.uint32)
            k = _I()
            self.api("b12x_continuation", self.h, self.scheme(grid, mp).h, _depth(depth), _ptr(p), p.shape[0],
                     _ptr(counts), cap, ctypes.byref(k))
            if k.value <= cap:
                return [int(c) for c in counts[: k.value]]
            cap = k.value  # a longer run than the buffer: run again with room for every count

    def searchByPoints(self, points, depth=-1):
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.dim)
        out = np.empty(p.shape[0], np.uint32)
        self.api("b12x_search_by_points", self.h, _ptr(p), p.shape[0], _depth(depth), _ptr(out))
        return out

    def subdivision_steps(self, grid, mp, steps, sd_flag=SD):
        """steps x {subdivide(SD); setFlagsByIterators(OverLeaves, HIT); remove(HIT); unsetFlags(HIT)}"""
        counts = np.zeros(4 * max(steps, 1), np.uint32)
        self.api("b12x_subdivision_steps", self.h, self.scheme(grid, mp).h, sd_flag, int(steps), _ptr(counts))
        return counts[: 4 * steps].reshape(steps, 4)

    def chain_recurrent_steps(self, grid, mp, steps, sd_flag=SD):
        """steps x {subdivide(SD); the leaves' matrix; strong components; HIT on the leaves of
        components with more than one leaf; remove(HIT); unsetFlags(HIT)} on the device
        (chainRecurrentSet.m:18-44). Per step: leaves after subdividing, components, components
        of more than one leaf, leaves kept."""
        counts = np.zeros(4 * max(steps, 1), np.uint32)
        self.api("b12x_chain_recurrent_steps", self.h, self.scheme(grid, mp).h, sd_flag, int(steps), _ptr(counts))
        return counts[: 4 * steps].reshape(steps, 4)

    # -- transition matrix
    def transitionMatrix(self, it, depth=-1, use_abs=False, row_major=False, use_invalid=False):
        h = _P()
        self.api("b12x_transition_matrix", self.h, it.scheme.h, it.depth, it.flags_to_check, _depth(depth),
                 int(use_abs), int(row_major), int(use_invalid), ctypes.byref(h))
        return CooMatrix(h, self.api)

    def transitionMatrixForLeaves(self, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverLeaves(grid, mp), LEAF, use_abs, row_major,
                                     use_invalid)

    def transitionMatrixForDepth(self, depth, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverDepth(depth, grid, mp), depth, use_abs, row_major,
                                     use_invalid)

    def transitionMatrixForRoot(self, grid, mp, use_abs=False, row_major=False, use_invalid=False):
        return self.transitionMatrix(self.makePointIteratorPairOverRoot(grid, mp), 0, use_abs, row_major, use_invalid)


def continuation(ibt, grid, mp, depths, points, on_step=None):
    """The continuation algorithm for unstable manifolds (B12c/src/matlab/continuation.m:1-63).

    For each depth d: reset the tree to the surrounding box (deleteDepth(0)), insert the boxes of
    depth d holding `points` with INS, then, while count(d) grows, turn INS into EXPD on depth d,
    insert the boxes of depth d hit by the images of the EXPD boxes with INS, and unset EXPD on
    the leaves. `points` is an (n, dim) array. Returns, per depth, the list of count(d) values:
    the first after the initial insertion, then one per expansion step (the last equals the one
    before it). `on_step(d, step, count)` is called after every expansion step. Any tree object
    with the same method names works, so the oracle's tree runs the same driver."""
    out = []
    if hasattr(ibt, "continuation_run") and on_step is None:  # the GPU tree: the loop on the device
        return [ibt.continuation_run(grid, mp, d, points) for d in depths]
    for d in depths:
        ibt.deleteDepth(0)
        ibt.insertByPoints(points, d, INS)
        nb0, nb1 = 0, ibt.count(d)
        it = ibt.makePointIteratorPairByFlags(d, EXPD, grid, mp)
        counts = [nb1]
        while nb1 > nb0:
            ibt.changeFlags(INS, EXPD, d)
            ibt.insertByIterators(it, d, INS, NONE)
            ibt.unsetFlags(EXPD, LEAF)
            nb0, nb1 = nb1, ibt.count(d)
            counts.append(nb1)
            if on_step is not None:
                on_step(d, len(counts) - 1, nb1)
        out.append(counts)
    return out


def fixedBoxes(ibt, grid, mp, steps=64):
    """Boxes that hit themselves (B12c/src/matlab/fixedBoxes.m:1-23): reset to the surrounding box,
    SD on every leaf, then `steps` times (the tree's maximal depth in the reference): subdivide,
    search the images of every leaf's test points, HIT on the leaves one of whose points lands in
    the leaf itself, remove(HIT), unsetFlags(HIT). Returns the boxes' centres and radii, one row
    per leaf (ibt.boxes without the flags). Any tree object with these method names works."""
    ibt.deleteDepth(0)
    ibt.setFlags(SD, LEAF)
    for _ in range(steps):
        ibt.subdivide(SD)
        it = ibt.makePointIteratorPairOverLeaves(grid, mp)
        s = ibt.searchByIterators(it, LEAF)
        n = ibt.size()
        own = s.reshape(n, -1) == np.arange(n, dtype=np.uint32)[:, None]
        ibt.setFlagsByStencil(own.any(axis=1).astype(np.uint8), HIT)
        ibt.remove(HIT)
        ibt.unsetFlags(HIT, LEAF)
    b = ibt.boxes(LEAF)
    return b[:, : b.shape[1] - 1]


def subdivisionByNim(ibt, grid, mp, steps, max_iter=1_000_000, tol=1e-12, threshold=1e-9):
    """The subdivision algorithm with selection by the natural invariant measure
    (B12c/src/matlab/subdivisionByNim.m:1-87): SD on every leaf, then per step subdivide(SD), the
    leaves' transition matrix, its invariant measure mu (the reference takes MATLAB's eigs or
    power method; here the power iteration mu <- mu P / ||mu P||_1 to ||mu_{k+1} - mu_k||_1 <= tol,
    the same measure as the measure step), then HIT on the leaves with mu >= threshold / n,
    remove(HIT), unsetFlags(HIT). Returns per step (leaves after subdividing, power iterations,
    leaves kept). Any tree object with these method names works."""
    out = []
    ibt.setFlags(SD, LEAF)
    for _ in range(steps):
        ibt.subdivide(SD)
        m = ibt.transitionMatrixForLeaves(grid, mp)
        mu, iters, _ = m.power_iteration(max_iter, tol)
        n = len(mu)
        v = np.abs(mu) / np.abs(mu).sum()
        ibt.setFlagsByStencil((v >= threshold / n).astype(np.uint8), HIT)
        ibt.remove(HIT)
        ibt.unsetFlags(HIT, LEAF)
        out.append((n, iters, ibt.count(LEAF)))
    return out


class KrylovBasis:
    """Device Krylov basis of a matrix (b12x_krylov_*): start, Arnoldi expansion with classical
    Gram-Schmidt twice, rotation by a small matrix, vector download."""

    def __init__(self, matrix, ncv):
        self.api, self.n, self.ncv = matrix.api, matrix.nRows, int(ncv)
        h = _P()
        self.api("b12x_krylov_create", matrix.h, self.ncv, ctypes.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.api.lib.b12x_krylov_destroy(self.h)
            self.h = None

    def start(self, v0=None):
        norm = _D()
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, np.float64)
            if v0.shape != (self.n,):
                raise ValueError(f"v0 must hold {self.n} values")
        self.api("b12x_krylov_start", self.h, _ptr(v0) if v0 is not None else None, ctypes.byref(norm))
        return norm.value

    def expand(self, p):
        h = np.zeros(p + 1, np.float64)
        beta = _D()
        self.api("b12x_krylov_expand", self.h, int(p), _ptr(h), ctypes.byref(beta))
        return h, beta.value

    def rotate(self, m, k, q):
        q = np.ascontiguousarray(q, np.float64)
        if q.shape != (m, k):
            raise ValueError("q must be m x k")
        self.api("b12x_krylov_rotate", self.h, int(m), int(k), _ptr(q))

    def vector(self, c):
        out = np.empty(self.n, np.float64)
        self.api("b12x_krylov_vector", self.h, int(c), _ptr(out))
        return out


def eigs(matrix, nev=4, ncv=None, tol=1e-10, max_restarts=1000, v0=None, want_vectors=False, on_restart=None):
    """The nev eigenvalues of largest modulus of a transition matrix by the Krylov-Schur method
    (Stewart 2002), as MATLAB's eigs (which subdivisionByNim.m:30-56 calls) computes them; the
    operator is the row-vector product v -> v P, so the eigenvectors are left eigenvectors.

    The basis and every product, dot product and update over the boxes live in the matrix's
    backend (matrix.krylov(ncv): the GPU's b12x_krylov_*, or the oracle's restatement); this
    driver keeps the small projected matrix. It holds A V_p = V_p B_p + v_p b_p^T (V orthonormal,
    A v = v P); an expansion step appends w = v_p P orthogonalised twice against V (h, beta):
    B_{p+1} = [[B_p, h[:p]], [b_p^T, h[p]]], b_{p+1} = beta e_p. At p = ncv:
     - the wanted set: the nev Ritz values of largest modulus (nev + 1 when the nev-th and the
       next are a complex pair). B = Z T Z^T (real Schur form, scipy.linalg.schur) with the wanted
       set first gives the partial Schur form A (V Z_w) = (V Z_w) T_ww + v_p (b^T Z)_w, whose
       residual ||(b^T Z)_w|| / min |lambda_wanted| is the stopping test (<= tol). It measures the
       wanted invariant subspace, so a cluster of nearly equal eigenvalues (Hénon-Heiles' symmetric
       almost-invariant pairs) converges with it; each pair's own residual |(b^T Z)_w . s| / |lambda|
       (s a unit eigenvector of T_ww) is reported beside it.
     - the restart: the same form with the k = nev + (ncv - nev) // 2 Ritz values of largest
       modulus first (one more when the cut would split a complex pair): V_k = V Z_k, v_k = v_ncv,
       B = T_kk, b = (b^T Z)_k.

    ncv defaults to 63 (at most n - 1), the most the basis holds: the Hénon-Heiles matrices'
    leading eigenvalues crowd together in modulus, and smaller bases converged to a wrong set (24
    vectors at depth 18 and 40 at depth 28 each missed a real eigenvalue of the leading four).

    Returns a dict: eigenvalues (by decreasing modulus), residual (the partial Schur form's),
    pair_residuals (the same order as the eigenvalues), restarts, matvecs (operator products),
    converged, ncv, and with want_vectors the Ritz vectors (one row each, complex).
    on_restart(restarts, matvecs, residual, eigenvalues), when given, is called at every test."""
    import scipy.linalg as sla

    n = matrix.nRows
    if ncv is None:
        ncv = min(63, n - 1)
    if not (1 <= nev < ncv <= n - 1):
        raise ValueError("need 1 <= nev < ncv <= n - 1")

    def leading(Bp, count):
        """Real Schur form of Bp with its `count` eigenvalues of largest modulus first (one more
        when the cut splits a complex pair): (T, Z, number first)."""
        mod = np.sort(np.abs(np.linalg.eigvals(Bp)))[::-1]
        p = len(mod)
        if count < p and mod[count - 1] == mod[count]:
            count += 1
        if count >= p:
            T, Z = sla.schur(Bp, output="real")
            return T, Z, p
        thr = 0.5 * (mod[count - 1] + mod[count])
        T, Z, sdim = sla.schur(Bp, output="real", sort=lambda re, im: re * re + im * im > thr * thr)
        return T, Z, sdim

    K = matrix.krylov(ncv)
    K.start(v0)
    B = np.zeros((ncv, ncv))
    b = np.zeros(ncv)
    p, matvecs, restarts = 0, 0, 0
    while True:
        breakdown = False
        while p < ncv:
            h, beta = K.expand(p)
            matvecs += 1
            B[:p, p] = h[:p]
            B[p, :p] = b[:p]
            B[p, p] = h[p]
            b[:] = 0.0
            b[p] = beta
            p += 1
            if beta == 0.0:  # the basis spans an invariant subspace: its Ritz values are exact
                breakdown = True
                break
        Bp, bp = B[:p, :p], b[:p]
        Tw, Zw, nw = leading(Bp, nev)
        rw = bp @ Zw
        lam, S = np.linalg.eig(Tw[:nw, :nw])
        order = np.argsort(-np.abs(lam), kind="stable")
        lam, S = lam[order], S[:, order]
        resid = float(np.linalg.norm(rw[:nw]) / max(np.min(np.abs(lam)), 1e-300))
        pair_res = np.array([abs(rw[:nw] @ S[:, i]) for i in range(nw)]) / np.maximum(np.abs(lam), 1e-300)
        converged = breakdown or resid <= tol
        if on_restart is not None:
            on_restart(restarts, matvecs, resid, lam)
        if converged or restarts >= max_restarts:
            out = {"eigenvalues": lam, "residual": resid, "pair_residuals": pair_res, "restarts": restarts,
                   "matvecs": matvecs, "converged": bool(converged), "ncv": ncv}
            if want_vectors:
                # Ritz vectors V Z_w s_i (real and imaginary parts rotated on the device)
                K.rotate(p, nw, Zw[:, :nw])
                basis = np.stack([K.vector(c) for c in range(nw)])
                out["vectors"] = S.T @ basis
            return out
        T, Z, keep = leading(Bp, min(nev + (ncv - nev) // 2, p))
        r = bp @ Z
        K.rotate(p, keep, Z[:, :keep])
        B = np.zeros((ncv, ncv))
        B[:keep, :keep] = T[:keep, :keep]
        b = np.zeros(ncv)
        b[:keep] = r[:keep]
        p = keep
        restarts += 1


def chainRecurrentSet(ibt, grid, mp, steps, on_step=None):
    """The chain recurrent set by subdivision (B12c/src/matlab/chainRecurrentSet.m:1-64): SD on
    every leaf, then per step subdivide(SD), the leaves' transition matrix, its strong
    components, HIT on the leaves whose component holds more than one leaf, remove(HIT),
    unsetFlags(HIT). Returns per step (leaves after subdividing, components, components of
    more than one leaf, leaves kept). Any tree object with the same method names works, so the
    oracle's tree runs the same driver; the GPU tree's chain_recurrent_steps runs the same steps
    without leaving the device."""
    out = []
    ibt.setFlags(SD, LEAF)
    for i in range(steps):
        ibt.subdivide(SD)
        m = ibt.transitionMatrixForLeaves(grid, mp)
        lab, n_comp, n_big = m.scc()
        size = np.bincount(lab, minlength=len(lab))
        n = len(lab)
        ibt.setFlagsByStencil(size[lab] > 1, HIT)
        ibt.remove(HIT)
        ibt.unsetFlags(HIT, LEAF)
        out.append((n, n_comp, n_big, ibt.count(LEAF)))
        if on_step is not None:
            on_step(i, out[-1])
    return out


class Forest:
    """Many subdivision jobs of one (grid, map kind) run in shared launches (batched sweeps and
    continuous-batching serving). Each job has its own domain, map parameters and depth."""

    def __init__(self, dim, grid, map_kind, use_abs=False, row_major=False, pi_max_iter=200, pi_tol=-1.0,
                 flavour="parity"):
        self.api = _Api(flavour)
        self.dim = dim
        gp = np.asarray(grid.params, np.uint64)
        h = _P()
        self.api("b12x_forest_create", dim, grid.kind, _ptr(gp), len(grid.params), map_kind, int(use_abs),
                 int(row_major), int(pi_max_iter), float(pi_tol), ctypes.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.api.lib.b12x_forest_destroy(self.h)
            self.h = None

    def add(self, center, radius, mp, depth):
        c = np.ascontiguousarray(center, np.float64)
        r = np.ascontiguousarray(radius, np.float64)
        p = np.ascontiguousarray(mp.params if mp.params else [0.0], np.float64)
        j = _U32()
        self.api("b12x_forest_add", self.h, _ptr(c), _ptr(r), _ptr(p), len(mp.params), int(depth), ctypes.byref(j))
        return j.value

    def step(self):
        a, r = _U32(), _U32()
        self.api("b12x_forest_step", self.h, ctypes.byref(a), ctypes.byref(r))
        return a.value, r.value

    def run(self):
        self.api("b12x_forest_run", self.h)

    def info(self, job):
        st, nl, nnz, it, res = _I(), _U32(), _U64(), _I(), _D()
        self.api("b12x_forest_job_info", self.h, job, ctypes.byref(st), ctypes.byref(nl), ctypes.byref(nnz),
                 ctypes.byref(it), ctypes.byref(res))
        return {"state": st.value, "n": nl.value, "nnz": nnz.value, "iters": it.value, "resid": res.value}

    def cover(self, job):
        n = self.info(job)["n"]
        b, d, f = np.empty(n, np.uint64), np.empty(n, np.uint8), np.empty(n, np.uint8)
        self.api("b12x_forest_job_cover", self.h, job, _ptr(b), _ptr(d), _ptr(f))
        return b, d, f

    def matrix(self, job):
        nnz = self.info(job)["nnz"]
        r, c, v = np.empty(nnz, np.uint32), np.empty(nnz, np.uint32), np.empty(nnz, np.float64)
        self.api("b12x_forest_job_matrix", self.h, job, _ptr(r), _ptr(c), _ptr(v))
        return r, c, v

    def measure(self, job):
        mu = np.empty(self.info(job)["n"], np.float64)
        self.api("b12x_forest_job_measure", self.h, job, _ptr(mu))
        return mu

    def release(self, job):
        self.api("b12x_forest_job_release", self.h, job)

    def footprint(self):
        b = _U64()
        self.api("b12x_forest_footprint", self.h, ctypes.byref(b))
        return b.value

    STAT_KEYS = ("steps", "subdiv_s", "retire_s", "pool_s", "pool_iters", "pool_max_jobs", "pool_max_cols",
                 "pool_max_nnz", "pool_col_iters", "pool_nnz_iters")

    def stats(self):
        v = np.zeros(len(self.STAT_KEYS), np.float64)
        self.api("b12x_forest_stats", self.h, _ptr(v), len(v))
        return dict(zip(self.STAT_KEYS, v.tolist()))

    def set_pool_iterations(self, iters):
        self.api("b12x_forest_set_pool_iterations", self.h, int(iters))


def serve_run(forests, arrival, forest_of, centers, radii, map_params, depths, budget_bytes=200e9,
              reserve_per_job=64e6):
    """Continuous-batching request loop (b12x_serve_run); returns (latency, completion, job_ids,
    info) with info = {admitted: per-request admission times, peak_footprint: bytes,
    max_in_flight, deferred: times the budget held a waiting request back}."""
    api = forests[0].api
    n = len(arrival)
    arr = (_P * len(forests))(*[f.h for f in forests])
    arrival = np.ascontiguousarray(arrival, np.float64)
    forest_of = np.ascontiguousarray(forest_of, np.int32)
    stride = max(f.dim for f in forests)
    C = np.zeros((n, stride))
    R = np.zeros((n, stride))
    for i in range(n):
        C[i, :len(centers[i])] = centers[i]
        R[i, :len(radii[i])] = radii[i]
    nmp = max(len(p) for p in map_params) if n else 0
    MP = np.zeros((n, max(nmp, 1)))
    for i in range(n):
        MP[i, :len(map_params[i])] = map_params[i]
    depths = np.ascontiguousarray(depths, np.uint8)
    lat = np.empty(n, np.float64)
    comp = np.empty(n, np.float64)
    ids = np.empty(n, np.uint32)
    adm = np.empty(n, np.float64)
    st = np.zeros(3, np.float64)
    api("b12x_serve_run", arr, len(forests), n, _ptr(arrival), _ptr(forest_of), _ptr(C), _ptr(R), stride, _ptr(MP),
        max(nmp, 1), _ptr(depths), float(budget_bytes), float(reserve_per_job), _ptr(lat), _ptr(comp), _ptr(ids),
        _ptr(adm), _ptr(st))
    return lat, comp, ids, {"admitted": adm, "peak_footprint": float(st[0]), "max_in_flight": int(st[1]),
                            "deferred": int(st[2])}
