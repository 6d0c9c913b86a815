// C entry points over the b12cpu oracle, for tests and for the CPU-baseline leg of bench.py.
// TEST INFRASTRUCTURE ONLY: the product library never links this file.
#include <cstdio>
#include <omp.h>
#include <memory>

#include "b12cpu.hpp"

using namespace b12cpu;

namespace {

struct SchemeBase {
  virtual ~SchemeBase() = default;
  int dim = 0;
};

template <int DIM>
struct Scheme : SchemeBase {
  Grid<DIM> grid;
  Map<DIM> map;
};

struct TreeBase {
  virtual ~TreeBase() = default;
  int dim = 0;
};

template <int DIM>
struct TreeH : TreeBase {
  Tree<DIM> t;
};

struct CooH {
  Coo m;
};

thread_local std::string g_err;

int fail(const std::string& s) {
  g_err = s;
  return -1;
}

template <int DIM>
SchemeBase* makeScheme(int gk, const uint64_t* gp, int ngp, int mk, const double* mp, int nmp) {
  auto* s = new Scheme<DIM>();
  s->dim = DIM;
  s->grid.kind = gk;
  if (gk == GRID_INNER) {
    s->grid.n = ngp > 0 ? gp[0] : 1;
    uint64_t k = 1;
    for (int j = 0; j < DIM; ++j) k *= s->grid.n;
    s->grid.k = k;
  } else {
    s->grid.k = ngp > 0 ? gp[0] : 1;
    s->grid.seed = ngp > 1 ? gp[1] : 0;
  }
  Map<DIM>& m = s->map;
  m.kind = mk;
  auto P = [&](int i, double d) { return i < nmp ? mp[i] : d; };
  switch (mk) {
    case MAP_IDENTITY: break;
    case MAP_HENON: m.p0 = P(0, 1.4); m.p1 = P(1, 0.3); m.steps = uint32_t(P(2, 1)); break;
    case MAP_LORENZ:
      m.p0 = P(0, 10.0); m.p1 = P(1, 28.0); m.p2 = P(2, 8.0 / 3.0); m.h = P(3, 0.01); m.steps = uint32_t(P(4, 20));
      break;
    case MAP_HENON_HEILES:
      m.order = int(P(0, 2)); m.h = P(1, 0.02); m.steps = uint32_t(P(2, 50));
      // cbrt(2) correctly rounded, the value g++ folds cbrtreal(2.0) to (MapTools.hpp:468)
      m.cbrt2 = 0x1.428a2f98d728bp+0; m.factor = 1.0 / (2.0 - m.cbrt2);
      break;
    case MAP_DUFFING: {
      m.p0 = P(0, 0.25); m.p1 = P(1, -1.0); m.p2 = P(2, 1.0);
      double gamma = P(3, 0.3), omega = P(4, 1.0), t0 = P(5, 0.0), t1 = P(6, 6.283185307179586),
             h = P(7, 6.283185307179586 / 100.0);
      m.forcing = buildForcingTable(t0, t1, h, gamma, omega);
      break;
    }
    default: delete s; return nullptr;
  }
  return s;
}

template <int DIM, class F>
int withTree(void* h, F f) {
  auto* t = static_cast<TreeBase*>(h);
  if (!t) return fail("null tree");
  if (t->dim != DIM) return fail("dim mismatch");
  return f(static_cast<TreeH<DIM>*>(t)->t);
}

#define DISPATCH(dim, CALL)                       \
  switch (dim) {                                  \
    case 1: { constexpr int D = 1; CALL; }        \
    case 2: { constexpr int D = 2; CALL; }        \
    case 3: { constexpr int D = 3; CALL; }        \
    case 4: { constexpr int D = 4; CALL; }        \
    default: return fail("dim must be 1..4");     \
  }

}  // namespace

extern "C" {

const char* b12cpu_last_error() { return g_err.c_str(); }

int b12cpu_scheme_create(int dim, int gridKind, const uint64_t* gp, int ngp, int mapKind, const double* mp,
                         int nmp, void** out) {
  SchemeBase* s = nullptr;
  switch (dim) {
    case 1: s = makeScheme<1>(gridKind, gp, ngp, mapKind, mp, nmp); break;
    case 2: s = makeScheme<2>(gridKind, gp, ngp, mapKind, mp, nmp); break;
    case 3: s = makeScheme<3>(gridKind, gp, ngp, mapKind, mp, nmp); break;
    case 4: s = makeScheme<4>(gridKind, gp, ngp, mapKind, mp, nmp); break;
    default: return fail("dim must be 1..4");
  }
  if (!s) return fail("bad map kind");
  *out = s;
  return 0;
}

void b12cpu_scheme_destroy(void* s) { delete static_cast<SchemeBase*>(s); }

uint64_t b12cpu_scheme_points_per_box(void* sh) {
  auto* s = static_cast<SchemeBase*>(sh);
  switch (s->dim) {
    case 1: return static_cast<Scheme<1>*>(s)->grid.k;
    case 2: return static_cast<Scheme<2>*>(s)->grid.k;
    case 3: return static_cast<Scheme<3>*>(s)->grid.k;
    case 4: return static_cast<Scheme<4>*>(s)->grid.k;
  }
  return 0;
}

// Apply the map alone to n points (row-major n x dim).
int b12cpu_map_points(void* sh, const double* in, double* out, uint64_t n) {
  auto* s = static_cast<SchemeBase*>(sh);
  DISPATCH(s->dim, {
    auto* sc = static_cast<Scheme<D>*>(s);
    for (uint64_t i = 0; i < n; ++i) sc->map(in + i * D, out + i * D);
    return 0;
  });
}

// Grid offsets g(i) for global point indices i0..i0+n-1 (row-major n x dim).
int b12cpu_grid_points(void* sh, uint64_t i0, uint64_t n, double* out) {
  auto* s = static_cast<SchemeBase*>(sh);
  DISPATCH(s->dim, {
    auto* sc = static_cast<Scheme<D>*>(s);
    for (uint64_t i = 0; i < n; ++i) sc->grid(i0 + i, out + i * D);
    return 0;
  });
}

int b12cpu_tree_create(int dim, const double* c, const double* r, const uint8_t* sd, int nsd, void** out) {
  TreeBase* t = nullptr;
  auto init = [&](auto* th) {
    for (int j = 0; j < th->dim; ++j) { th->t.geo.center[j] = c[j]; th->t.geo.radius[j] = r[j]; }
    if (sd && !th->t.geo.setScheme(sd, nsd)) return false;
    return true;
  };
  bool ok = false;
  switch (dim) {
    case 1: { auto* x = new TreeH<1>(); x->dim = 1; ok = init(x); t = x; break; }
    case 2: { auto* x = new TreeH<2>(); x->dim = 2; ok = init(x); t = x; break; }
    case 3: { auto* x = new TreeH<3>(); x->dim = 3; ok = init(x); t = x; break; }
    case 4: { auto* x = new TreeH<4>(); x->dim = 4; ok = init(x); t = x; break; }
    default: return fail("dim must be 1..4");
  }
  if (!ok) { delete t; return fail("sdScheme has values >= dim"); }
  *out = t;
  return 0;
}

void b12cpu_tree_destroy(void* h) { delete static_cast<TreeBase*>(h); }

#define TREE_CALL(h, BODY)                                                 \
  try {                                                                    \
    auto* tb = static_cast<TreeBase*>(h);                                  \
    if (!tb) return fail("null tree");                                     \
    DISPATCH(tb->dim, { auto& T = static_cast<TreeH<D>*>(tb)->t; BODY; }); \
  } catch (const std::exception& e) {                                      \
    return fail(e.what());                                                 \
  }

// OpenMP threads of the oracle's parallel loops (the CPU baseline runs with all of the process's
// cores and with one, as B12's OMP and CPP host backends do); returns the previous count
int b12cpu_set_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}

int b12cpu_set_options(void* h, int evalTwice, uint64_t chunkPoints) {
  TREE_CALL(h, { T.evalTwice = evalTwice != 0; T.chunkPoints = chunkPoints ? chunkPoints : ~0ull; return 0; });
}

int b12cpu_size(void* h, uint32_t* out) { TREE_CALL(h, { *out = T.size(); return 0; }); }
int b12cpu_depth(void* h, uint8_t* out) { TREE_CALL(h, { *out = T.depth(); return 0; }); }
int b12cpu_count(void* h, uint8_t d, uint32_t* out) { TREE_CALL(h, { *out = T.count(d); return 0; }); }

// boxes(depth): *n boxes of 2*dim+1 doubles each; out may be null to ask for *n only
int b12cpu_boxes(void* h, uint8_t d, double* out, uint64_t* n) {
  TREE_CALL(h, {
    const std::vector<double> b = T.boxes(d);
    *n = b.size() / (2 * D + 1);
    if (out) std::copy(b.begin(), b.end(), out);
    return 0;
  });
}

// enumerate(): *n = depth() + 1 counts; out may be null
int b12cpu_enumerate(void* h, uint32_t* out, int* n) {
  TREE_CALL(h, {
    const std::vector<NrBoxes> e = T.enumerate();
    *n = int(e.size());
    if (out) std::copy(e.begin(), e.end(), out);
    return 0;
  });
}

int b12cpu_download(void* h, uint64_t* bits, uint8_t* depth, uint8_t* flags) {
  TREE_CALL(h, {
    std::memcpy(bits, T.bits.data(), T.bits.size() * 8);
    std::memcpy(depth, T.depthV.data(), T.depthV.size());
    std::memcpy(flags, T.flagsV.data(), T.flagsV.size());
    return 0;
  });
}

// Leaves are put into B12's order (isStrictlyPreceding, helpFunctions.hpp:40-98) as every
// B12 tree is kept sorted.
int b12cpu_upload(void* h, uint32_t n, const uint64_t* bits, const uint8_t* depth, const uint8_t* flags) {
  TREE_CALL(h, {
    T.resize(n);
    std::memcpy(T.bits.data(), bits, size_t(n) * 8);
    std::memcpy(T.depthV.data(), depth, n);
    std::memcpy(T.flagsV.data(), flags, n);
    T.sortBoxes();
    T.insertSurroundingBoxIfEmpty();
    return 0;
  });
}

int b12cpu_subdivide(void* h, uint8_t f, uint32_t* ret) { TREE_CALL(h, { *ret = T.subdivide(f); return 0; }); }
int b12cpu_unsubdivide(void* h, uint8_t f, uint32_t* ret) { TREE_CALL(h, { *ret = T.unsubdivide(f); return 0; }); }
int b12cpu_remove(void* h, uint8_t f, uint32_t* ret) { TREE_CALL(h, { *ret = T.remove(f); return 0; }); }
int b12cpu_set_flags(void* h, uint8_t f, uint8_t d, uint32_t* ret) {
  TREE_CALL(h, { *ret = T.setFlags(f, d); return 0; });
}
int b12cpu_unset_flags(void* h, uint8_t f, uint8_t d, uint32_t* ret) {
  TREE_CALL(h, { *ret = T.unsetFlags(f, d); return 0; });
}
int b12cpu_change_flags(void* h, uint8_t fu, uint8_t fs, uint8_t d, uint32_t* ret) {
  TREE_CALL(h, { *ret = T.changeFlags(fu, fs, d); return 0; });
}
// op: 0 set, 1 unset, 2 change (fs used only for change)
int b12cpu_flags_by_stencil(void* h, int op, const uint8_t* st, uint8_t f, uint8_t fs, uint8_t d, uint32_t* ret) {
  TREE_CALL(h, {
    *ret = op == 0 ? T.setFlagsByStencil(st, f, d)
                   : (op == 1 ? T.unsetFlagsByStencil(st, f, d) : T.changeFlagsByStencil(st, f, fs, d));
    return 0;
  });
}

uint64_t b12cpu_num_points(void* h, void* sh, uint8_t dIt, uint8_t check) {
  try {
    auto* tb = static_cast<TreeBase*>(h);
    switch (tb->dim) {
      case 1: { auto* s = static_cast<Scheme<1>*>(sh); return static_cast<TreeH<1>*>(tb)->t.makePoints(s->grid, s->map, dIt, check).size(); }
      case 2: { auto* s = static_cast<Scheme<2>*>(sh); return static_cast<TreeH<2>*>(tb)->t.makePoints(s->grid, s->map, dIt, check).size(); }
      case 3: { auto* s = static_cast<Scheme<3>*>(sh); return static_cast<TreeH<3>*>(tb)->t.makePoints(s->grid, s->map, dIt, check).size(); }
      case 4: { auto* s = static_cast<Scheme<4>*>(sh); return static_cast<TreeH<4>*>(tb)->t.makePoints(s->grid, s->map, dIt, check).size(); }
    }
  } catch (...) {
  }
  return 0;
}

// op: 0 set, 1 unset, 2 change flags of boxes on depth d hit by the image points
int b12cpu_flags_by_image(void* h, void* sh, uint8_t dIt, uint8_t check, int op, uint8_t f, uint8_t fs, uint8_t d,
                          uint32_t* ret) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    auto sv = T.searchCompacted(ps, d);
    *ret = op == 0 ? T.setFlagsBySearch(sv, f, d)
                   : (op == 1 ? T.unsetFlagsBySearch(sv, f, d) : T.changeFlagsBySearch(sv, f, fs, d));
    return 0;
  });
}

int b12cpu_search_by_image(void* h, void* sh, uint8_t dIt, uint8_t check, uint8_t d, uint32_t* out) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    auto v = T.search(ps, d);
    std::memcpy(out, v.data(), v.size() * 4);
    return 0;
  });
}

int b12cpu_search_compacted(void* h, void* sh, uint8_t dIt, uint8_t check, uint8_t d, uint32_t* out,
                            uint64_t* nOut) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    auto v = T.searchCompacted(ps, d);
    if (out) std::memcpy(out, v.data(), v.size() * 4);
    *nOut = v.size();
    return 0;
  });
}

// Max-depth keys of the image points (bits, depth) in enumeration order.
int b12cpu_point_boxes(void* h, void* sh, uint8_t dIt, uint8_t check, uint64_t* bits, uint8_t* depth) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    for (uint64_t i = 0; i < ps.size(); ++i) {
      Box b = T.pointBox(ps, i);
      bits[i] = b.bits;
      depth[i] = b.depth;
    }
    return 0;
  });
}

// ImplicitBoxTree.hpp:1215-1262 deleteDepth
int b12cpu_delete_depth(void* h, uint8_t d, uint32_t* ret) { TREE_CALL(h, { *ret = T.deleteDepth(d); return 0; }); }

// ImplicitBoxTree.hpp:1613-1892 insertByIterators(points, depth, flagsForInserted, flagsForHit):
// res (may be null) receives isToInsert per point
int b12cpu_insert_by_image(void* h, void* sh, uint8_t dIt, uint8_t check, uint8_t d, uint8_t fIns, uint8_t fHit,
                           uint8_t* res) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    auto v = T.insertByIterators(ps, d, fIns, fHit);
    if (res && !v.empty()) std::memcpy(res, v.data(), v.size());
    return 0;
  });
}

// ImplicitBoxTree.hpp:1894-1930 insertByPoints
int b12cpu_insert_by_points(void* h, const double* pts, uint64_t n, uint8_t d, uint8_t fIns, uint8_t fHit,
                            uint8_t* res) {
  TREE_CALL(h, {
    auto v = T.insertByPoints(pts, n, d, fIns, fHit);
    if (res && !v.empty()) std::memcpy(res, v.data(), v.size());
    return 0;
  });
}

int b12cpu_search_by_points(void* h, const double* pts, uint64_t n, uint8_t d, uint32_t* out) {
  TREE_CALL(h, {
    auto v = T.searchPoints(pts, n, d);
    std::memcpy(out, v.data(), v.size() * 4);
    return 0;
  });
}

int b12cpu_transition_matrix(void* h, void* sh, uint8_t dIt, uint8_t check, uint8_t d, int useAbs, int rowMajor,
                             int useInvalid, void** out) {
  TREE_CALL(h, {
    auto* s = static_cast<Scheme<D>*>(sh);
    auto ps = T.makePoints(s->grid, s->map, dIt, check);
    auto* c = new CooH();
    c->m = T.transitionMatrix(ps, d, useAbs != 0, rowMajor != 0, useInvalid != 0);
    *out = c;
    return 0;
  });
}

uint64_t b12cpu_coo_nnz(void* c) { return static_cast<CooH*>(c)->m.rows.size(); }

int b12cpu_coo_get(void* c, uint32_t* rows, uint32_t* cols, double* vals, uint32_t* nRows, uint32_t* nCols) {
  const Coo& m = static_cast<CooH*>(c)->m;
  if (rows) std::memcpy(rows, m.rows.data(), m.rows.size() * 4);
  if (cols) std::memcpy(cols, m.cols.data(), m.cols.size() * 4);
  if (vals) std::memcpy(vals, m.vals.data(), m.vals.size() * 8);
  *nRows = m.nRows;
  *nCols = m.nCols;
  return 0;
}

int b12cpu_coo_create(const uint32_t* rows, const uint32_t* cols, const double* vals, uint64_t nnz, uint32_t nRows,
                      uint32_t nCols, void** out) {
  auto* c = new CooH();
  c->m.rows.assign(rows, rows + nnz);
  c->m.cols.assign(cols, cols + nnz);
  c->m.vals.assign(vals, vals + nnz);
  c->m.nRows = nRows;
  c->m.nCols = nCols;
  *out = c;
  return 0;
}

void b12cpu_coo_destroy(void* c) { delete static_cast<CooH*>(c); }

int b12cpu_compress_index_vector(const uint32_t* idx, uint64_t n, uint32_t dim, uint32_t* out) {
  std::vector<NrBoxes> v(idx, idx + n);
  auto o = compressIndexVector(v, dim);
  std::memcpy(out, o.data(), o.size() * 4);
  return 0;
}

// mu: in/out, length nRows (pass all zeros to start from the uniform vector)
int b12cpu_power_iteration(void* c, int maxIter, double tol, double* mu, int* iters, double* resid) {
  const Coo& m = static_cast<CooH*>(c)->m;
  std::vector<double> v(mu, mu + m.nRows);
  bool zero = true;
  for (double x : v) zero = zero && x == 0.0;
  if (zero) v.clear();
  *iters = powerIteration(m, maxIter, tol, v, resid);
  std::memcpy(mu, v.data(), v.size() * 8);
  return 0;
}

// stencil of the leaves meeting the Hénon-Heiles energy band (4-dimensional trees only)
int b12cpu_energy_band_stencil(void* h, double emin, double emax, uint8_t* out) {
  auto* tb = static_cast<TreeBase*>(h);
  if (tb->dim != 4) return fail("energy band: 4-dimensional trees only");
  auto st = energyBandStencil(static_cast<TreeH<4>*>(tb)->t, emin, emax);
  std::memcpy(out, st.data(), st.size());
  return 0;
}

// H: m x m Ritz matrix (row-major); X: m x nRows basis (column by column), may be null; the
// residual is that of the leading `lead` basis columns
int b12cpu_subspace_iteration(void* c, int m, int lead, int maxIter, double tol, double* H, double* X, int* iters,
                              double* resid) {
  const Coo& P = static_cast<CooH*>(c)->m;
  if (m < 1 || m > 8 || lead < 1 || lead > m || maxIter < 0)
    return fail("subspace iteration: 1 <= lead <= m <= 8, max_iter >= 0");
  std::vector<double> h, x;
  double r = -1.0;
  const int k = subspaceIteration(P, m, lead, maxIter, tol, h, x, &r);
  if (k < 0) return fail("subspace iteration: Gram matrix not positive definite");
  std::memcpy(H, h.data(), h.size() * 8);
  if (X) std::memcpy(X, x.data(), x.size() * 8);
  if (iters) *iters = k;
  if (resid) *resid = r;
  return 0;
}

// Krylov basis (KrylovBasis in b12cpu.hpp): the same operations as b12x_krylov_*
int b12cpu_krylov_create(void* c, int ncv, void** out) {
  const Coo& P = static_cast<CooH*>(c)->m;
  if (ncv < 2 || ncv + 1 > 64 || P.nRows < NrBoxes(ncv) + 1) return fail("krylov: 2 <= ncv <= 63, ncv < nRows");
  *out = new KrylovBasis(P, ncv);
  return 0;
}
int b12cpu_krylov_destroy(void* k) {
  delete static_cast<KrylovBasis*>(k);
  return 0;
}
int b12cpu_krylov_start(void* k, const double* v0, double* norm) {
  const double b = static_cast<KrylovBasis*>(k)->start(v0);
  if (!(b > 0.0)) return fail("krylov: zero start vector");
  if (norm) *norm = b;
  return 0;
}
int b12cpu_krylov_expand(void* k, int p, double* h, double* beta) {
  auto* K = static_cast<KrylovBasis*>(k);
  if (p < 0 || p >= K->ncv) return fail("krylov: p must lie in [0, ncv)");
  *beta = K->expand(p, h);
  return 0;
}
int b12cpu_krylov_rotate(void* k, int m, int kk, const double* q) {
  auto* K = static_cast<KrylovBasis*>(k);
  if (m < 1 || m > K->ncv || kk < 1 || kk > m) return fail("krylov: need 1 <= k <= m <= ncv");
  K->rotate(m, kk, q);
  return 0;
}
int b12cpu_krylov_vector(void* k, int c, double* out) {
  auto* K = static_cast<KrylovBasis*>(k);
  if (c < 0 || c > K->ncv) return fail("krylov: column out of range");
  std::copy(K->col(c), K->col(c) + K->n, out);
  return 0;
}

// labels (nRows entries): the largest vertex index of the vertex's strong component; counts of
// components and of components with more than one vertex
int b12cpu_coo_scc(void* c, uint32_t* labels, uint32_t* nComp, uint32_t* nBig) {
  const Coo& P = static_cast<CooH*>(c)->m;
  const std::vector<NrBoxes> lab = strongComponents(P);
  std::vector<uint32_t> size(P.nRows, 0);
  for (NrBoxes l : lab) ++size[l];
  uint32_t comps = 0, big = 0;
  for (NrBoxes v = 0; v < P.nRows; ++v)
    if (lab[v] == v) {
      ++comps;
      big += size[v] > 1 ? 1 : 0;
    }
  if (labels) std::memcpy(labels, lab.data(), lab.size() * 4);
  if (nComp) *nComp = comps;
  if (nBig) *nBig = big;
  return 0;
}

}  // extern "C"
