// C++ façade of B12's tree / matrix / functor API over the b12x C ABI (b12x_c_api.h).
//
// A B12 program switches its hot path to the MI355X by including this header instead of B12's
// ImplicitBoxTree.h, Grids.h and Map_*.h and using the architecture tag b12::MI355X:
//
//   ImplicitBoxTree<MI355X, 2, double, 64> ibt(center, radius);      // ImplicitBoxTree.h:49-51
//   auto itPair = ibt.makePointIteratorPairOverLeaves(grid, map);    // ImplicitBoxTree.h:308-326
//   ibt.setFlagsByIterators(itPair, Flag::HIT, -1);                  // ImplicitBoxTree.h:468-469
//
// Names, argument meaning and return values follow the reference (citations relative to
// B12c/src/include/). ThrustSystem<MI355X>::Vector<T> is a host std::vector: the tree's own
// arrays live in HBM behind the handle, and vectors passed in or returned are host copies, as in
// B12's MEX binding. Error convention as in B12: invalid input prints "Warning: ..." to
// std::cout and leaves the tree unchanged (ImplicitBoxTree.hpp:3501,3548); a GPU runtime failure
// throws std::runtime_error (B12 surfaces those as thrust::system_error).
#pragma once

#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../b12x_c_api.h"

namespace b12 {

enum Architecture { MI355X = 16 };

using Depth = uint8_t;       // TypeDefinitions.h:19
using Dimension = uint8_t;   // TypeDefinitions.h:37
using NrBoxes = uint32_t;    // TypeDefinitions.h:45
using NrPoints = uint64_t;   // TypeDefinitions.h:57
using Flags = uint8_t;       // Flags.h:9

enum Flag : Flags { NONE = 0x00, HIT = 0x01, INS = 0x02, EXPD = 0x04, SD = 0x08, EQU = 0x10, ENTR = 0x20,
                    EXTR = 0x40, ALL = 0xFF };  // Flags.h:11-22

inline Depth getLeafDepth() { return Depth(-1); }          // TypeDefinitions.hpp
inline NrBoxes getInvalidNrBoxes() { return NrBoxes(-1); }  // TypeDefinitions.hpp

template <Architecture A>
struct ThrustSystem {
  template <typename T>
  using Vector = std::vector<T>;
};

namespace detail {
inline bool check(int rc, const char* what) {
  if (rc == B12X_OK) return true;
  if (rc == B12X_E_HIP || rc == B12X_E_NOMEM || rc == B12X_E_NODEVICE)
    throw std::runtime_error(std::string(what) + ": " + b12x_last_error());
  std::cout << "Warning: " << b12x_last_error() << std::endl;
  return false;
}

struct SchemeHolder {
  b12x_scheme h = nullptr;
  ~SchemeHolder() {
    if (h) b12x_scheme_destroy(h);
  }
};
}  // namespace detail

// ---- Grids (Grids.h:307-370): kind + parameters ------------------------------------------------
template <Dimension DIM, typename REAL>
struct InnerGrid {
  NrPoints nPointsPerDim_;
  explicit InnerGrid(NrPoints nPointsPerDim = 1) : nPointsPerDim_(nPointsPerDim) {}
  NrPoints getNumberOfPointsPerBox() const {
    NrPoints k = 1;
    for (int j = 0; j < DIM; ++j) k *= nPointsPerDim_;
    return k;
  }
  Dimension dim() const { return DIM; }
  int kind() const { return B12X_GRID_INNER; }
  std::vector<uint64_t> params() const { return {nPointsPerDim_}; }
};

template <Dimension DIM, typename REAL>
struct UniformGrid {
  NrPoints nPoints_;
  unsigned long long seed_;
  explicit UniformGrid(NrPoints nPoints = 1, unsigned long long seed = 0) : nPoints_(nPoints), seed_(seed) {}
  NrPoints getNumberOfPointsPerBox() const { return nPoints_; }
  Dimension dim() const { return DIM; }
  int kind() const { return B12X_GRID_UNIFORM; }
  std::vector<uint64_t> params() const { return {nPoints_, seed_}; }
};

// ---- Maps (Map_Henon.h:26-40, Map_Lorenz.h:29-44, Map_Identity.h:9-21) ---------------------------
template <Dimension DIM, typename REAL>
struct Henon {
  REAL a_, b_;
  uint32_t steps_;
  explicit Henon(REAL a = 1.4, REAL b = 0.3, uint32_t steps = 1) : a_(a), b_(b), steps_(steps) {}
  int kind() const { return B12X_MAP_HENON; }
  std::vector<double> params() const { return {double(a_), double(b_), double(steps_)}; }
};

template <Dimension DIM, typename REAL>
struct Lorenz {
  REAL sigma_, rho_, beta_, h_;
  uint32_t steps_;
  explicit Lorenz(REAL sigma = 10.0, REAL rho = 28.0, REAL beta = 8.0 / 3.0, REAL h = 0.01, uint32_t steps = 20)
      : sigma_(sigma), rho_(rho), beta_(beta), h_(h), steps_(steps) {}
  int kind() const { return B12X_MAP_LORENZ; }
  std::vector<double> params() const {
    return {double(sigma_), double(rho_), double(beta_), double(h_), double(steps_)};
  }
};

template <Dimension DIM, typename REAL>
struct Identity {
  int kind() const { return B12X_MAP_IDENTITY; }
  std::vector<double> params() const { return {}; }
};

// Hénon–Heiles flow as SymplecticIntegratorAuto<4> (MapTools.hpp:361-523); no B12 counterpart map.
template <Dimension DIM, typename REAL>
struct HenonHeiles {
  uint8_t order_;
  REAL h_;
  uint32_t steps_;
  explicit HenonHeiles(uint8_t order = 2, REAL h = 0.02, uint32_t steps = 50) : order_(order), h_(h), steps_(steps) {}
  int kind() const { return B12X_MAP_HENON_HEILES; }
  std::vector<double> params() const { return {double(order_), double(h_), double(steps_)}; }
};

// Forced Duffing period map as RK4 (MapTools.hpp:35-108); no B12 counterpart map.
template <Dimension DIM, typename REAL>
struct ForcedDuffing {
  REAL delta_, alpha_, beta_, gamma_, omega_, t0_, tEnd_, h_;
  explicit ForcedDuffing(REAL delta = 0.25, REAL alpha = -1.0, REAL beta = 1.0, REAL gamma = 0.3, REAL omega = 1.0,
                         REAL t0 = 0.0, REAL tEnd = 6.283185307179586, REAL h = 6.283185307179586 / 100.0)
      : delta_(delta), alpha_(alpha), beta_(beta), gamma_(gamma), omega_(omega), t0_(t0), tEnd_(tEnd), h_(h) {}
  int kind() const { return B12X_MAP_DUFFING; }
  std::vector<double> params() const { return {delta_, alpha_, beta_, gamma_, omega_, t0_, tEnd_, h_}; }
};

// The lazy point range of makePointIteratorPair* (ImplicitBoxTree.h:246-430): instead of a
// thrust iterator pair it carries the compiled (grid, map) scheme and the iterator's depth/flags.
struct PointIteratorPair {
  std::shared_ptr<detail::SchemeHolder> scheme;
  Depth depth;
  Flags flagsToCheck;
  NrPoints nPointsPerBox;
};

// ---- CooMatrix (CooMatrix.h:56-121) -----------------------------------------------------------
template <Architecture A, typename T>
struct CooMatrix {
  typename ThrustSystem<A>::template Vector<NrBoxes> rowIndices_, columnIndices_;
  typename ThrustSystem<A>::template Vector<T> values_;
  NrBoxes nRows_ = 0, nColumns_ = 0;

  CooMatrix() = default;

  // CooMatrix.h:75-76: the matrix saved in fileName (or fileName + ".json")
  explicit CooMatrix(const std::string& fileName) {
    b12x_coo h = nullptr;
    if (!detail::check(b12x_coo_load_json(fileName.c_str(), &h), "CooMatrix")) return;
    uint64_t nnz = 0;
    uint32_t nr = 0, nc = 0;
    int rm = 0;
    b12x_coo_info(h, &nnz, &nr, &nc, &rm);
    rowIndices_.resize(nnz);
    columnIndices_.resize(nnz);
    std::vector<double> v(nnz);
    if (nnz) b12x_coo_download(h, rowIndices_.data(), columnIndices_.data(), v.data());
    values_.assign(v.begin(), v.end());
    nRows_ = nr;
    nColumns_ = nc;
    b12x_coo_destroy(h);
  }

  // CooMatrix.h:61-91, .hpp:304-447: ordering and comparison, on the device
  bool isSortedRowMajor() const { return sorted(1); }
  bool isSortedColumnMajor() const { return sorted(0); }
  CooMatrix& sortRowMajor() {
    sortBy(1);
    return *this;
  }
  CooMatrix& sortColumnMajor() {
    sortBy(0);
    return *this;
  }
  // CooMatrix.hpp:376-426: the symmetric (or antisymmetric) part, in place
  void makeSymmetric(bool makeAntisymmetric = false) {
    b12x_coo h = toDevice();
    if (!h) return;
    uint64_t nnz = 0;
    if (detail::check(b12x_coo_make_symmetric(h, makeAntisymmetric ? 1 : 0), "makeSymmetric") &&
        detail::check(b12x_coo_info(h, &nnz, nullptr, nullptr, nullptr), "makeSymmetric")) {
      std::vector<double> v(nnz);
      rowIndices_.resize(nnz);
      columnIndices_.resize(nnz);
      if (nnz) detail::check(b12x_coo_download(h, rowIndices_.data(), columnIndices_.data(), v.data()), "makeSymmetric");
      values_.assign(v.begin(), v.end());
    }
    b12x_coo_destroy(h);
  }
  // the same sizes and entries independent of the orderings (CooMatrix.hpp:304-352); this
  // matrix's entries are compared with the first entries of the other, as the reference does
  template <Architecture A2, typename T2>
  bool operator==(const CooMatrix<A2, T2>& o) const {
    if (nRows_ != o.nRows_ || nColumns_ != o.nColumns_) return false;
    const bool frm = isSortedRowMajor(), fcm = !frm && isSortedColumnMajor();
    const bool srm = o.isSortedRowMajor(), scm = !srm && o.isSortedColumnMajor();
    if (frm && !srm) return *this == CooMatrix<A2, T2>(o).sortRowMajor();
    if (fcm && !scm) return *this == CooMatrix<A2, T2>(o).sortColumnMajor();
    if (srm && !frm) return CooMatrix(*this).sortRowMajor() == o;
    if (scm && !fcm) return CooMatrix(*this).sortColumnMajor() == o;
    bool res = o.values_.size() >= values_.size();
    for (size_t e = 0; res && e < values_.size(); ++e)
      res = rowIndices_[e] == o.rowIndices_[e] && columnIndices_[e] == o.columnIndices_[e] &&
            double(values_[e]) == double(o.values_[e]);
    if (!res && !((frm && srm) || (fcm && scm)))
      res = CooMatrix(*this).sortRowMajor() == CooMatrix<A2, T2>(o).sortRowMajor();
    return res;
  }

  // CooMatrix.h:110-112: B12's JSON file (".json" appended when missing)
  void save(const std::string& fileName) const {
    b12x_coo h = nullptr;
    std::vector<double> v(values_.begin(), values_.end());
    if (!detail::check(b12x_coo_upload(rowIndices_.data(), columnIndices_.data(), v.data(), rowIndices_.size(),
                                       nRows_, nColumns_, &h),
                       "save"))
      return;
    detail::check(b12x_coo_save_json(h, fileName.c_str()), "save");
    b12x_coo_destroy(h);
  }
  // the matrix on the device for one operation (uploads a copy of the host vectors)
  b12x_coo toDevice() const {
    b12x_coo h = nullptr;
    std::vector<double> v(values_.begin(), values_.end());
    detail::check(b12x_coo_upload(rowIndices_.data(), columnIndices_.data(), v.data(), rowIndices_.size(), nRows_,
                                  nColumns_, &h),
                  "CooMatrix");
    return h;
  }
  bool sorted(int rowMajor) const {
    b12x_coo h = toDevice();
    int r = 0;
    if (h) detail::check(b12x_coo_is_sorted(h, rowMajor, &r), "isSorted");
    b12x_coo_destroy(h);
    return r != 0;
  }
  void sortBy(int rowMajor) {
    b12x_coo h = toDevice();
    if (!h) return;
    std::vector<double> v(values_.size());
    if (detail::check(b12x_coo_sort(h, rowMajor), "sort") && !v.empty() &&
        detail::check(b12x_coo_download(h, rowIndices_.data(), columnIndices_.data(), v.data()), "sort"))
      values_.assign(v.begin(), v.end());
    b12x_coo_destroy(h);
  }
};

// compressIndexVector<A>(indices, dim) — CooMatrix.h:20-23
template <Architecture A>
inline void compressIndexVector(typename ThrustSystem<A>::template Vector<NrBoxes>& indices, NrBoxes dim) {
  std::vector<NrBoxes> out(size_t(dim) + 1);
  if (detail::check(b12x_compress_index_vector(indices.data(), indices.size(), dim, out.data()),
                    "compressIndexVector"))
    indices.swap(out);
}

// Invariant measure of a row-stochastic transition matrix (build-defined; B12 does this step in
// MATLAB, subdivisionByNim.m:30-59): returns mu with mu = mu P / ||mu P||_1.
template <Architecture A, typename T>
inline std::vector<double> invariantMeasure(const CooMatrix<A, T>& m, int maxIter, double tol, int* iters = nullptr,
                                            double* resid = nullptr) {
  b12x_coo h = nullptr;
  std::vector<double> vals(m.values_.begin(), m.values_.end());
  detail::check(b12x_coo_upload(m.rowIndices_.data(), m.columnIndices_.data(), vals.data(), m.rowIndices_.size(),
                                m.nRows_, m.nColumns_, &h),
                "invariantMeasure");
  std::vector<double> mu(m.nRows_);
  int it = 0;
  double r = 0.0;
  detail::check(b12x_power_iteration(h, maxIter, tol, 0, mu.data(), &it, &r), "invariantMeasure");
  b12x_coo_destroy(h);
  if (iters) *iters = it;
  if (resid) *resid = r;
  return mu;
}

// ---- ImplicitBoxTree (ImplicitBoxTree.h:32-640) ---------------------------------------------------
template <Architecture A, Dimension DIM, typename REAL = double, Depth N = 64>
class ImplicitBoxTree {
  static_assert(A == MI355X, "this header implements the MI355X architecture");
  static_assert(N == 64, "bit patterns of 64 bits (every configuration on the path fits N = 64)");
  static_assert(DIM >= 1 && DIM <= 4, "dimensions 1..4");

 public:
  template <typename T>
  using Vector = typename ThrustSystem<A>::template Vector<T>;

  // ImplicitBoxTree.h:44 default constructor: the unit hypercube [0,1]^DIM
  ImplicitBoxTree() : ImplicitBoxTree(Vector<REAL>(DIM, REAL(0.5)), Vector<REAL>(DIM, REAL(0.5))) {}

  // ImplicitBoxTree.h:49-51
  template <typename REAL2>
  ImplicitBoxTree(const Vector<REAL2>& center, const Vector<REAL2>& radius) {
    create(center, radius, nullptr, 0);
  }

  // ImplicitBoxTree.h:53-56
  template <typename REAL2, typename UInt>
  ImplicitBoxTree(const Vector<REAL2>& center, const Vector<REAL2>& radius, const Vector<UInt>& sdScheme) {
    std::vector<uint8_t> sd(sdScheme.begin(), sdScheme.end());
    create(center, radius, sd.data(), int(sd.size()));
  }

  // ImplicitBoxTree.h:58-59: the tree saved in fileName (a JSON file, fileName + ".json", or a
  // legacy binary file); on failure a warning and the default tree, as the reference does
  explicit ImplicitBoxTree(const std::string& fileName) {
    if (!detail::check(b12x_tree_load(fileName.c_str(), DIM, &h_), "ImplicitBoxTree")) {
      h_ = nullptr;
      create(Vector<REAL>(DIM, REAL(0.5)), Vector<REAL>(DIM, REAL(0.5)), nullptr, 0);
    }
  }

  // ImplicitBoxTree.h:70-76
  void save(const std::string& fileName) const { detail::check(b12x_tree_save_json(h_, fileName.c_str()), "save"); }
  void saveForGaio(const std::string& fileName) const {
    detail::check(b12x_tree_save_gaio(h_, fileName.c_str()), "saveForGaio");
  }

  // ImplicitBoxTree.h:46-47, 64-67: copies of the domain, scheme, leaves and host buffer size
  ImplicitBoxTree(const ImplicitBoxTree& o) { copyFrom(o); }
  ImplicitBoxTree& operator=(const ImplicitBoxTree& o) {
    if (this != &o) {
      if (h_) b12x_tree_destroy(h_);
      h_ = nullptr;
      copyFrom(o);
    }
    return *this;
  }
  ~ImplicitBoxTree() {
    if (h_) b12x_tree_destroy(h_);
  }

  // ---- domain and bookkeeping (ImplicitBoxTree.h:82-85, 141-143, 244, 550-569)
  Vector<REAL> getCenter() const { return geometry(0); }
  Vector<REAL> getRadius() const { return geometry(1); }
  Vector<Dimension> getSdScheme() const {
    int dim = 0, nsd = 0;
    uint8_t sd[65];
    detail::check(b12x_tree_geometry(h_, &dim, nullptr, nullptr, sd, &nsd), "getSdScheme");
    return Vector<Dimension>(sd, sd + nsd);
  }
  // setCenter / setRadius: cut to DIM entries or extended with 0.5 (ImplicitBoxTree.hpp:3470-3505)
  template <typename REAL2>
  Vector<REAL> setCenter(const Vector<REAL2>& center) {
    double c[DIM];
    for (int j = 0; j < DIM; ++j) c[j] = j < int(center.size()) ? double(center[j]) : 0.5;
    detail::check(b12x_tree_set_domain(h_, c, nullptr, nullptr, 0, 0), "setCenter");
    return getCenter();
  }
  template <typename REAL2>
  Vector<REAL> setRadius(const Vector<REAL2>& radius) {
    double r[DIM];
    for (int j = 0; j < DIM; ++j) r[j] = j < int(radius.size()) ? double(radius[j]) : 0.5;
    for (const auto& v : radius)
      if (v < REAL2(0)) {
        std::cout << "Warning: Input has negative values. Radius vector remains unchanged." << std::endl;
        return getRadius();
      }
    detail::check(b12x_tree_set_domain(h_, nullptr, r, nullptr, 0, 0), "setRadius");
    return getRadius();
  }
  // setSdScheme: repeated periodically, the default scheme when empty (ImplicitBoxTree.hpp:3513-3552)
  template <typename UInt>
  Vector<Dimension> setSdScheme(const Vector<UInt>& sdScheme) {
    std::vector<uint8_t> sd;
    for (const auto& v : sdScheme) {
      if (!(v >= 0 && v < UInt(DIM))) {
        std::cout << "Warning: Input has values >= DIM. SdScheme remains unchanged." << std::endl;
        return getSdScheme();
      }
      sd.push_back(uint8_t(v));
    }
    detail::check(b12x_tree_set_domain(h_, nullptr, nullptr, sd.empty() ? nullptr : sd.data(), int(sd.size()), 1),
                  "setSdScheme");
    return getSdScheme();
  }
  bool isSdSchemeDefault() const {
    const Vector<Dimension> sd = getSdScheme();
    for (size_t i = 0; i < sd.size(); ++i)
      if (sd[i] != Dimension(i % DIM)) return false;
    return true;
  }
  // sdCount_ (adaptSdCount, ImplicitBoxTree.hpp:574-587): (64 + 2) x DIM bisection counts
  Vector<Depth> getSdCount() const {
    const Vector<Dimension> sd = getSdScheme();
    Vector<Depth> c(size_t(64 + 2) * DIM, 0);
    for (size_t i = 0; i < sd.size(); ++i) {
      for (int j = 0; j < DIM; ++j) c[(i + 1) * DIM + j] = c[i * DIM + j];
      ++c[(i + 1) * DIM + sd[i]];
    }
    return c;
  }
  bool areAllLeavesOnDepth(Depth depth) const {
    int r = 0;
    detail::check(b12x_all_leaves_on_depth(h_, depth, &r), "areAllLeavesOnDepth");
    return r != 0;
  }
  template <Architecture A2, Dimension DIM2, typename REAL2, Depth N2>
  bool hasSameBoxes(const ImplicitBoxTree<A2, DIM2, REAL2, N2>& other) const {
    int r = 0;
    detail::check(b12x_has_same_boxes(h_, other.handle(), &r), "hasSameBoxes");
    return r != 0;
  }
  uint64_t getNHostBufferBytes() const {
    uint64_t v = 0;
    detail::check(b12x_get_host_buffer_bytes(h_, &v), "getNHostBufferBytes");
    return v;
  }
  uint64_t setNHostBufferBytes(uint64_t n) {
    detail::check(b12x_set_host_buffer_bytes(h_, n), "setNHostBufferBytes");
    return getNHostBufferBytes();
  }

  // ---- queries
  NrBoxes size() const { return ret([&](uint32_t* r) { return b12x_size(h_, r); }, "size"); }
  NrBoxes count(Depth depth = getLeafDepth()) const {
    return ret([&](uint32_t* r) { return b12x_count(h_, depth, r); }, "count");
  }
  Depth depth() const {
    uint8_t d = 0;
    detail::check(b12x_depth(h_, &d), "depth");
    return d;
  }
  Dimension dim() const { return DIM; }
  // boxes(depth) — ImplicitBoxTree.h:86-91: {centre, radius, flags} of every box at depth
  typename ThrustSystem<A>::template Vector<REAL> boxes(Depth depth) const {
    uint64_t n = 0;
    detail::check(b12x_boxes(h_, depth, nullptr, &n), "boxes");
    typename ThrustSystem<A>::template Vector<REAL> res(n * (2 * DIM + 1));
    if (n) detail::check(b12x_boxes(h_, depth, res.data(), &n), "boxes");
    return res;
  }
  void boxes(Depth depth, REAL* res_begin) const {
    uint64_t n = 0;
    detail::check(b12x_boxes(h_, depth, res_begin, &n), "boxes");
  }

  // enumerate() — ImplicitBoxTree.h:132-133: {count(0), ..., count(depth())}
  typename ThrustSystem<A>::template Vector<NrBoxes> enumerate() const {
    int n = 0;
    detail::check(b12x_enumerate(h_, nullptr, &n), "enumerate");
    typename ThrustSystem<A>::template Vector<NrBoxes> res(n);
    detail::check(b12x_enumerate(h_, res.data(), &n), "enumerate");
    return res;
  }

  // bitPatternVector_, depthVector_, flagsVector_ as host copies (ImplicitBoxTree.h:619-624)
  void getLeaves(Vector<uint64_t>& bits, Vector<Depth>& depths, Vector<Flags>& flags) const {
    const NrBoxes n = size();
    bits.resize(n);
    depths.resize(n);
    flags.resize(n);
    detail::check(b12x_tree_download(h_, bits.data(), depths.data(), flags.data()), "getLeaves");
  }
  Vector<Depth> getDepthVector() const {
    Vector<uint64_t> b;
    Vector<Depth> d;
    Vector<Flags> f;
    getLeaves(b, d, f);
    return d;
  }
  Vector<Flags> getFlagsVector() const {
    Vector<uint64_t> b;
    Vector<Depth> d;
    Vector<Flags> f;
    getLeaves(b, d, f);
    return f;
  }

  // ---- structure
  NrBoxes subdivide(Flags flags = Flag::SD) { return ret([&](uint32_t* r) { return b12x_subdivide(h_, flags, r); }, "subdivide"); }
  NrBoxes unsubdivide(Flags flags = Flag::SD) {
    return ret([&](uint32_t* r) { return b12x_unsubdivide(h_, flags, r); }, "unsubdivide");
  }
  NrBoxes remove(Flags flags) { return ret([&](uint32_t* r) { return b12x_remove(h_, flags, r); }, "remove"); }
  // deleteDepth(depth) — ImplicitBoxTree.h:124-127
  NrBoxes deleteDepth(Depth depth) {
    return ret([&](uint32_t* r) { return b12x_delete_depth(h_, depth, r); }, "deleteDepth");
  }

  // ---- insertion (ImplicitBoxTree.h:170-202); the result holds, per point, whether a box was
  // created for it
  Vector<bool> insertByIterators(const PointIteratorPair& points, Depth depth, Flags flagsForInserted = Flag::INS,
                                 Flags flagsForHit = Flag::NONE) {
    uint64_t n = 0;
    if (!detail::check(b12x_num_points(h_, points.scheme->h, points.depth, points.flagsToCheck, &n),
                       "insertByIterators"))
      return {};
    std::vector<uint8_t> r(n);
    detail::check(b12x_insert_by_image(h_, points.scheme->h, points.depth, points.flagsToCheck, depth,
                                       flagsForInserted, flagsForHit, r.data()),
                  "insertByIterators");
    return Vector<bool>(r.begin(), r.end());
  }
  Vector<bool> insertByPoints(const Vector<REAL>& points, Depth depth, Flags flagsForInserted = Flag::INS,
                              Flags flagsForHit = Flag::NONE) {
    std::vector<double> p(points.begin(), points.end());
    std::vector<uint8_t> r(p.size() / DIM);
    detail::check(b12x_insert_by_points(h_, p.data(), r.size(), depth, flagsForInserted, flagsForHit, r.data()),
                  "insertByPoints");
    return Vector<bool>(r.begin(), r.end());
  }

  // ---- flags (ImplicitBoxTree.h:93-108, 461-485, 527-541)
  NrBoxes setFlags(Flags flags, Depth depth = getLeafDepth()) {
    return ret([&](uint32_t* r) { return b12x_set_flags(h_, flags, depth, r); }, "setFlags");
  }
  NrBoxes unsetFlags(Flags flags, Depth depth = getLeafDepth()) {
    return ret([&](uint32_t* r) { return b12x_unset_flags(h_, flags, depth, r); }, "unsetFlags");
  }
  NrBoxes changeFlags(Flags flagsToUnset, Flags flagsToSet, Depth depth = getLeafDepth()) {
    return ret([&](uint32_t* r) { return b12x_change_flags(h_, flagsToUnset, flagsToSet, depth, r); }, "changeFlags");
  }
  template <typename BoolIterator>
  NrBoxes setFlagsByStencil(BoolIterator stencil_begin, Flags flags, Depth depth = getLeafDepth()) {
    return stencilOp(0, stencil_begin, flags, 0, depth);
  }
  template <typename BoolIterator>
  NrBoxes unsetFlagsByStencil(BoolIterator stencil_begin, Flags flags, Depth depth = getLeafDepth()) {
    return stencilOp(1, stencil_begin, flags, 0, depth);
  }
  template <typename BoolIterator>
  NrBoxes changeFlagsByStencil(BoolIterator stencil_begin, Flags flagsToUnset, Flags flagsToSet,
                               Depth depth = getLeafDepth()) {
    return stencilOp(2, stencil_begin, flagsToUnset, flagsToSet, depth);
  }
  NrBoxes setFlagsByIterators(const PointIteratorPair& points, Flags flags, Depth depth = getLeafDepth()) {
    return imageOp(points, 0, flags, 0, depth);
  }
  NrBoxes unsetFlagsByIterators(const PointIteratorPair& points, Flags flags, Depth depth = getLeafDepth()) {
    return imageOp(points, 1, flags, 0, depth);
  }
  NrBoxes changeFlagsByIterators(const PointIteratorPair& points, Flags flagsToUnset, Flags flagsToSet,
                                 Depth depth = getLeafDepth()) {
    return imageOp(points, 2, flagsToUnset, flagsToSet, depth);
  }
  // setFlagsBySearch(sorted unique searchVector, flags, depth) — ImplicitBoxTree.h:477-480
  NrBoxes setFlagsBySearch(const Vector<NrBoxes>& searchVector, Flags flags, Depth depth = getLeafDepth()) {
    return stencilOp(0, searchStencil(searchVector, depth).begin(), flags, 0, depth);
  }
  NrBoxes unsetFlagsBySearch(const Vector<NrBoxes>& searchVector, Flags flags, Depth depth = getLeafDepth()) {
    return stencilOp(1, searchStencil(searchVector, depth).begin(), flags, 0, depth);
  }
  NrBoxes changeFlagsBySearch(const Vector<NrBoxes>& searchVector, Flags flagsToUnset, Flags flagsToSet,
                              Depth depth = getLeafDepth()) {
    return stencilOp(2, searchStencil(searchVector, depth).begin(), flagsToUnset, flagsToSet, depth);
  }
  // setFlagsByPoints(points {x0..x_{DIM-1}; ...}, flags, depth) — ImplicitBoxTree.h:474-475, and the
  // unset / change forms (ImplicitBoxTree.h:100-101, 534-535)
  NrBoxes setFlagsByPoints(const Vector<REAL>& points, Flags flags, Depth depth = getLeafDepth()) {
    Vector<NrBoxes> s = searchByPoints(points, depth);
    return stencilOp(0, searchStencil(s, depth).begin(), flags, 0, depth);
  }
  NrBoxes unsetFlagsByPoints(const Vector<REAL>& points, Flags flags, Depth depth = getLeafDepth()) {
    Vector<NrBoxes> s = searchByPoints(points, depth);
    return stencilOp(1, searchStencil(s, depth).begin(), flags, 0, depth);
  }
  NrBoxes changeFlagsByPoints(const Vector<REAL>& points, Flags flagsToUnset, Flags flagsToSet,
                              Depth depth = getLeafDepth()) {
    Vector<NrBoxes> s = searchByPoints(points, depth);
    return stencilOp(2, searchStencil(s, depth).begin(), flagsToUnset, flagsToSet, depth);
  }

  // ---- point iterators (ImplicitBoxTree.h:266-430)
  template <typename Grid, typename Map>
  PointIteratorPair makePointIteratorPairOverLeaves(Grid grid, Map map, Flags flagsToSet = Flag::NONE) {
    return makePair(grid, map, getLeafDepth(), Flag::NONE);
  }
  template <typename Grid, typename Map>
  PointIteratorPair makePointIteratorPairOverRoot(Grid grid, Map map, Flags flagsToSet = Flag::NONE) {
    return makePair(grid, map, 0, Flag::NONE);
  }
  template <typename Grid, typename Map>
  PointIteratorPair makePointIteratorPairOverDepth(Depth depth, Grid grid, Map map, Flags flagsToSet = Flag::NONE) {
    return makePair(grid, map, depth, Flag::NONE);
  }
  template <typename Grid, typename Map>
  PointIteratorPair makePointIteratorPairByFlags(Depth depth, Flags flagsToCheck, Grid grid, Map map,
                                                 Flags flagsToSet = Flag::NONE) {
    return makePair(grid, map, depth, flagsToCheck);
  }

  // ---- search (ImplicitBoxTree.h:436-459)
  Vector<NrBoxes> searchByIterators(const PointIteratorPair& points, Depth depth = getLeafDepth()) {
    uint64_t n = 0;
    if (!detail::check(b12x_num_points(h_, points.scheme->h, points.depth, points.flagsToCheck, &n),
                       "searchByIterators"))
      return {};
    Vector<NrBoxes> out(n);
    detail::check(b12x_search_by_image(h_, points.scheme->h, points.depth, points.flagsToCheck, depth, out.data()),
                  "searchByIterators");
    return out;
  }
  Vector<NrBoxes> searchCompacted(const PointIteratorPair& points, Depth depth = getLeafDepth()) {
    Vector<NrBoxes> out(std::max<NrBoxes>(size(), 1));
    uint64_t n = 0;
    detail::check(b12x_search_compacted(h_, points.scheme->h, points.depth, points.flagsToCheck, depth, out.data(), &n),
                  "searchCompacted");
    out.resize(n);
    return out;
  }
  Vector<NrBoxes> searchByPoints(const Vector<REAL>& points, Depth depth = getLeafDepth()) {
    std::vector<double> p(points.begin(), points.end());
    Vector<NrBoxes> out(p.size() / DIM);
    detail::check(b12x_search_by_points(h_, p.data(), out.size(), depth, out.data()), "searchByPoints");
    return out;
  }

  // ---- transition matrices (ImplicitBoxTree.h:491-525)
  template <typename T>
  void transitionMatrix(const PointIteratorPair& points, Depth depth, NrPoints nPointsPerBox, bool useAbsoluteValues,
                        bool useRowMajor, bool useInvalidNrBoxes, CooMatrix<A, T>& res) {
    (void)nPointsPerBox;  // implied by the scheme's grid
    b12x_coo m = nullptr;
    if (!detail::check(b12x_transition_matrix(h_, points.scheme->h, points.depth, points.flagsToCheck, depth,
                                              useAbsoluteValues, useRowMajor, useInvalidNrBoxes, &m),
                       "transitionMatrix"))
      return;
    uint64_t nnz = 0;
    uint32_t nr = 0, nc = 0;
    int rm = 0;
    b12x_coo_info(m, &nnz, &nr, &nc, &rm);
    res.rowIndices_.resize(nnz);
    res.columnIndices_.resize(nnz);
    std::vector<double> v(nnz);
    detail::check(b12x_coo_download(m, res.rowIndices_.data(), res.columnIndices_.data(), v.data()),
                  "transitionMatrix");
    res.values_.assign(v.begin(), v.end());
    res.nRows_ = nr;
    res.nColumns_ = nc;
    b12x_coo_destroy(m);
  }
  template <typename T, typename Grid, typename Map>
  CooMatrix<A, T> transitionMatrixForLeaves(Grid grid, Map map, bool useAbsoluteValues = false,
                                            bool useRowMajor = false, bool useInvalidNrBoxes = false) {
    CooMatrix<A, T> res;
    transitionMatrix<T>(makePointIteratorPairOverLeaves(grid, map), getLeafDepth(), grid.getNumberOfPointsPerBox(),
                        useAbsoluteValues, useRowMajor, useInvalidNrBoxes, res);
    return res;
  }
  template <typename T, typename Grid, typename Map>
  CooMatrix<A, T> transitionMatrixForDepth(Depth depth, Grid grid, Map map, bool useAbsoluteValues = false,
                                           bool useRowMajor = false, bool useInvalidNrBoxes = false) {
    CooMatrix<A, T> res;
    transitionMatrix<T>(makePointIteratorPairOverDepth(depth, grid, map), depth, grid.getNumberOfPointsPerBox(),
                        useAbsoluteValues, useRowMajor, useInvalidNrBoxes, res);
    return res;
  }
  template <typename T, typename Grid, typename Map>
  CooMatrix<A, T> transitionMatrixForRoot(Grid grid, Map map, bool useAbsoluteValues = false,
                                          bool useRowMajor = false, bool useInvalidNrBoxes = false) {
    CooMatrix<A, T> res;
    transitionMatrix<T>(makePointIteratorPairOverRoot(grid, map), 0, grid.getNumberOfPointsPerBox(),
                        useAbsoluteValues, useRowMajor, useInvalidNrBoxes, res);
    return res;
  }

  // ---- fused subdivision steps (subdivision.m:18-25): returns the four counts per step
  template <typename Grid, typename Map>
  std::vector<NrBoxes> subdivisionSteps(Grid grid, Map map, int steps, Flags sd = Flag::SD) {
    PointIteratorPair p = makePair(grid, map, getLeafDepth(), Flag::NONE);
    std::vector<NrBoxes> counts(size_t(4) * std::max(steps, 1));
    detail::check(b12x_subdivision_steps(h_, p.scheme->h, sd, steps, counts.data()), "subdivisionSteps");
    counts.resize(size_t(4) * steps);
    return counts;
  }

  b12x_tree handle() const { return h_; }

 private:
  b12x_tree h_ = nullptr;

  template <typename REAL2>
  void create(const Vector<REAL2>& center, const Vector<REAL2>& radius, const uint8_t* sd, int nsd) {
    if (center.size() != DIM || radius.size() != DIM)
      throw std::invalid_argument("center and radius must have DIM entries");
    double c[DIM], r[DIM];
    for (int j = 0; j < DIM; ++j) {
      c[j] = double(center[j]);
      r[j] = double(radius[j]);
    }
    int rc = b12x_tree_create(DIM, c, r, sd, nsd, &h_);
    if (rc == B12X_E_ARG && sd) {  // ImplicitBoxTree.hpp:3548: warn and keep the default scheme
      std::cout << "Warning: Input has values >= DIM. SdScheme remains unchanged." << std::endl;
      rc = b12x_tree_create(DIM, c, r, nullptr, 0, &h_);
    }
    if (rc != B12X_OK) throw std::runtime_error(std::string("ImplicitBoxTree: ") + b12x_last_error());
  }

  template <typename F>
  NrBoxes ret(F f, const char* what) const {
    uint32_t r = 0;
    detail::check(f(&r), what);
    return r;
  }

  template <typename Grid, typename Map>
  PointIteratorPair makePair(const Grid& grid, const Map& map, Depth depth, Flags flagsToCheck) {
    auto holder = std::make_shared<detail::SchemeHolder>();
    std::vector<uint64_t> gp = grid.params();
    std::vector<double> mp = map.params();
    int rc = b12x_scheme_create(DIM, grid.kind(), gp.data(), int(gp.size()), map.kind(), mp.data(), int(mp.size()),
                                &holder->h);
    if (rc != B12X_OK) throw std::runtime_error(std::string("makePointIteratorPair: ") + b12x_last_error());
    return PointIteratorPair{holder, depth, flagsToCheck, grid.getNumberOfPointsPerBox()};
  }

  Vector<REAL> geometry(int which) const {
    int dim = 0;
    double c[4], r[4];
    detail::check(b12x_tree_geometry(h_, &dim, c, r, nullptr, nullptr), "geometry");
    const double* v = which == 0 ? c : r;
    return Vector<REAL>(v, v + DIM);
  }

  void copyFrom(const ImplicitBoxTree& o) {
    const Vector<Dimension> sd = o.getSdScheme();
    std::vector<uint8_t> s(sd.begin(), sd.end());
    const bool dflt = s.size() == 64;  // the default scheme holds 64 entries, a given one 65
    create(o.getCenter(), o.getRadius(), dflt ? nullptr : s.data(), dflt ? 0 : int(s.size()));
    std::vector<uint64_t> bits;
    std::vector<Depth> depths;
    std::vector<Flags> flags;
    o.getLeaves(bits, depths, flags);
    detail::check(b12x_tree_upload(h_, uint32_t(bits.size()), bits.data(), depths.data(), flags.data()), "copy");
    setNHostBufferBytes(o.getNHostBufferBytes());
  }

  // a stencil holds one entry per box of depth `depth` (count(depth))
  template <typename BoolIterator>
  NrBoxes stencilOp(int op, BoolIterator st, Flags a, Flags b, Depth depth) {
    const NrBoxes n = count(depth);
    std::vector<uint8_t> s(n);
    for (NrBoxes i = 0; i < n; ++i, ++st) s[i] = *st ? 1 : 0;
    return ret([&](uint32_t* r) { return b12x_flags_by_stencil(h_, op, s.data(), a, b, depth, r); }, "flagsByStencil");
  }

  std::vector<uint8_t> searchStencil(const Vector<NrBoxes>& sv, Depth depth) {
    std::vector<uint8_t> s(count(depth), 0);
    for (NrBoxes i : sv)
      if (i != getInvalidNrBoxes() && i < s.size()) s[i] = 1;
    return s;
  }

  NrBoxes imageOp(const PointIteratorPair& p, int op, Flags a, Flags b, Depth depth) {
    return ret([&](uint32_t* r) {
      return b12x_flags_by_image(h_, p.scheme->h, p.depth, p.flagsToCheck, op, a, b, depth, r);
    }, "flagsByIterators");
  }
};

}  // namespace b12
