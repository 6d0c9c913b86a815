// B12's documented C++ example (doc/B12_doc.tex:237-294: 20 subdivision steps of the Hénon map)
// compiled against the b12x façade: only the include and the architecture tag differ. Prints the
// leaf count per step, then the transition matrix size and the leading measure entries.
#include <cstdio>
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
  return 0;
}
