// B12's continuation algorithm (B12c/src/matlab/continuation.m) written against the b12x façade:
// the unstable manifold of the Hénon map's fixed point, one run per depth. Prints, per depth,
// "d: c0 c1 ... ck" with count(d) after the initial insertion and after every expansion step.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "b12x/ImplicitBoxTree.h"

using namespace b12;

int main(int argc, char** argv) {
  const Architecture A = MI355X;
  const Dimension DIM = 2;
  ThrustSystem<A>::Vector<double> center(DIM, 0.0);
  ThrustSystem<A>::Vector<double> radius(DIM, 3.0);
  ImplicitBoxTree<A, DIM, double, 64> ibt(center, radius);
  InnerGrid<DIM, double> grid(4);
  Henon<DIM, double> map(1.4, 0.3, 1);
  const double a = 1.4, b = 0.3;
  const double x = (-(1 - b) + std::sqrt((1 - b) * (1 - b) + 4 * a)) / (2 * a);
  ThrustSystem<A>::Vector<double> points{x, b * x};

  for (int i = 1; i < argc; ++i) {
    const Depth d = Depth(std::atoi(argv[i]));
    ibt.deleteDepth(0);
    ibt.insertByPoints(points, d, Flag::INS);
    NrBoxes nb0 = 0, nb1 = ibt.count(d);
    std::printf("%d: %u", int(d), nb1);
    auto scheme = ibt.makePointIteratorPairByFlags(d, Flag::EXPD, grid, map);
    while (nb1 > nb0) {
      ibt.changeFlags(Flag::INS, Flag::EXPD, d);
      ibt.insertByIterators(scheme, d, Flag::INS, Flag::NONE);
      ibt.unsetFlags(Flag::EXPD);
      nb0 = nb1;
      nb1 = ibt.count(d);
      std::printf(" %u", nb1);
    }
    std::printf("\n");
  }
  return 0;
}
