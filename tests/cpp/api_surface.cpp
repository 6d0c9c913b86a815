// The façade's domain, bookkeeping, copy and by-points/by-search methods (ImplicitBoxTree.h:
// 46-67, 82-108, 141-143, 244, 396-430, 472-480, 532-569), printed one result per line for the
// GPU test to compare with the Python mirror.
#include <cstdio>
#include <vector>

#include "b12x/ImplicitBoxTree.h"

using namespace b12;

template <typename V>
static void put(const char* name, const V& v) {
  std::printf("%s", name);
  for (const auto& x : v) std::printf(" %.17g", double(x));
  std::printf("\n");
}

int main() {
  const Architecture A = MI355X;
  ThrustSystem<A>::Vector<double> center{0.0, 0.0}, radius{1.5, 0.5};
  ImplicitBoxTree<A, 2, double, 64> t(center, radius);
  InnerGrid<2, double> grid(4);
  Henon<2, double> map(1.4, 0.3, 1);
  t.setFlags(Flag::SD, 0);
  t.subdivisionSteps(grid, map, 10);
  put("center", t.getCenter());
  put("radius", t.getRadius());
  std::printf("default %d\n", int(t.isSdSchemeDefault()));
  std::printf("alldepth %d %d %d\n", int(t.areAllLeavesOnDepth(getLeafDepth())), int(t.areAllLeavesOnDepth(10)),
              int(t.areAllLeavesOnDepth(254)));
  ImplicitBoxTree<A, 2, double, 64> u(t);  // copy
  std::printf("same %d\n", int(t.hasSameBoxes(u)));
  ThrustSystem<A>::Vector<double> pts{0.1, 0.05, -0.7, 0.2, 0.9, -0.3, 5.0, 5.0, 0.12, 0.051};
  std::printf("setByPoints %u\n", t.setFlagsByPoints(pts, 0x10));
  std::printf("changeByPoints %u\n", t.changeFlagsByPoints(pts, 0x10, 0x20, 6));
  std::printf("unsetByPoints %u\n", t.unsetFlagsByPoints(pts, 0x20, 6));
  std::printf("same %d\n", int(t.hasSameBoxes(u)));
  auto it = t.makePointIteratorPairOverDepth(7, grid, map);
  auto s = t.searchCompacted(it, 7);
  std::printf("changeBySearch %u\n", t.changeFlagsBySearch(s, 0x00, 0x40, 7));
  std::printf("unsetBySearch %u\n", t.unsetFlagsBySearch(s, 0x40, 7));
  auto m = t.transitionMatrixForDepth<double>(7, grid, map);
  std::printf("tm7 %zu %u %u\n", m.values_.size(), m.nRows_, m.nColumns_);
  std::printf("hostbytes %llu", (unsigned long long)t.getNHostBufferBytes());
  std::printf(" %llu\n", (unsigned long long)t.setNHostBufferBytes(uint64_t(1) << 20));
  ThrustSystem<A>::Vector<int> sd{1, 0, 0};
  put("sd", t.setSdScheme(sd));
  put("sdcount", t.getSdCount());
  std::printf("default %d\n", int(t.isSdSchemeDefault()));
  ThrustSystem<A>::Vector<double> bad{-1.0, 2.0};
  put("radius", t.setRadius(bad));  // rejected with a warning
  ThrustSystem<A>::Vector<double> c3{0.25};
  put("center", t.setCenter(c3));  // extended with 0.5
  std::printf("same %d\n", int(t.hasSameBoxes(u)));
  ImplicitBoxTree<A, 2, double, 64> w;
  w = t;  // assignment
  std::printf("same %d %llu\n", int(w.hasSameBoxes(t)), (unsigned long long)w.getNHostBufferBytes());
  // CooMatrix ordering and comparison
  CooMatrix<A, double> P = u.transitionMatrixForLeaves<double>(grid, map);            // column-major
  CooMatrix<A, double> Q = u.transitionMatrixForLeaves<double>(grid, map, false, true);  // row-major
  std::printf("coo %d %d %d %d %d\n", int(P.isSortedColumnMajor()), int(P.isSortedRowMajor()),
              int(Q.isSortedRowMajor()), int(P == Q), int(Q == P));
  CooMatrix<A, double> R(P);
  R.sortRowMajor();
  std::printf("sorted %d %d %d\n", int(R.isSortedRowMajor()), int(R.rowIndices_ == Q.rowIndices_ &&
              R.columnIndices_ == Q.columnIndices_ && R.values_ == Q.values_), int(R == P));
  R.values_[0] += 1.0;
  std::printf("differ %d\n", int(R == P));
  CooMatrix<A, double> S(Q);
  S.makeSymmetric();
  std::printf("symmetric %zu %d\n", S.values_.size(), int(S.isSortedRowMajor()));
  return 0;
}
