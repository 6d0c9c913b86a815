// B12's documented C++ example (doc/B12_doc.tex:237-294: 20 subdivision steps of the Hénon map)
// compiled against the b12x façade: only the include and the architecture tag differ. Prints the
// leaf count per step, then the transition matrix size and the leading measure entries. With a
// directory as second argument it also saves the tree and the matrix there and reads them back
// (an addition to the documented example).
#include <cstdio>
#include <string>
#include <iostream>

#include "b12x/ImplicitBoxTree.h"

using namespace b12;

int main(int argc, char** argv) {
  const Architecture A = MI355X;
  const Dimension DIM = 2;
  const int steps = argc > 1 ? std::atoi(argv[1]) : 20;

  ThrustSystem<A>::Vector<double> center(DIM, 0.0);
  ThrustSystem<A>::Vector<double> radius(DIM, 3.0);
  ImplicitBoxTree<A, DIM, double, 64> ibt(center, radius);
  InnerGrid<DIM, double> grid(10);
  Henon<DIM, double> map(1.4, 0.3, 1);

  ibt.setFlags(Flag::SD, 0);
  for (int i = 1; i <= steps; ++i) {
    ibt.subdivide(Flag::SD);
    auto itPair = ibt.makePointIteratorPairOverLeaves(grid, map);
    ibt.setFlagsByIterators(itPair, Flag::HIT, -1);
    ibt.remove(Flag::HIT);
    ibt.unsetFlags(Flag::HIT, -1);
    std::cout << i << ": " << ibt.count(-1) << std::endl;
  }
  CooMatrix<A, double> P = ibt.transitionMatrixForLeaves<double>(grid, map);
  std::cout << "nnz " << P.values_.size() << " rows " << P.nRows_ << " cols " << P.nColumns_ << std::endl;
  int iters = 0;
  double resid = 0;
  std::vector<double> mu = invariantMeasure(P, 50, -1.0, &iters, &resid);
  std::printf("mu0 %.17g iters %d\n", mu.empty() ? 0.0 : mu[0], iters);
  if (argc > 2) {  // files: ImplicitBoxTree::save and ImplicitBoxTree(fileName), the same for CooMatrix
    const std::string dir = argv[2];
    ibt.save(dir + "/tree");
    ImplicitBoxTree<A, DIM, double, 64> loaded(dir + "/tree");
    P.save(dir + "/matrix");
    CooMatrix<A, double> Q(dir + "/matrix");
    std::cout << "reloaded " << loaded.count(-1) << " " << Q.values_.size() << " "
              << (Q.rowIndices_ == P.rowIndices_ && Q.columnIndices_ == P.columnIndices_ && Q.values_ == P.values_)
              << std::endl;
  }
  return 0;
}
