// SymplecticIntegratorAuto is not built here: MapTools.hpp calls cbrtreal from mathFunctions.h,
// which includes curand_kernel.h (absent from this image).
// Driver that exposes the reference's own map functors (compiled from the reference's headers
// where they lie) through a C ABI, so the oracle's integrator restatement can be checked
// bit-for-bit against them. TEST INFRASTRUCTURE ONLY; the output goes to oracle/_ref/.
#include <cstdint>
#include <cmath>

#include "Map_Henon.h"
#include "Map_Identity.h"
#include "Map_Lorenz.h"

namespace {
// Forced Duffing right-hand side for the reference's non-autonomous RK4.
struct DuffRhs {
  double delta, alpha, beta, gamma, omega;
  __host__ __device__ void operator()(double t, const double* x, double* r) const {
    r[0] = x[1];
    r[1] = gamma * cos(omega * t) - delta * x[1] - alpha * x[0] - beta * x[0] * x[0] * x[0];
  }
};
}  // namespace

extern "C" {

int ref_henon(double a, double b, uint32_t steps, const double* in, double* out, uint64_t n) {
  b12::Henon<2, double> m(a, b, steps);
  for (uint64_t i = 0; i < n; ++i) m(in + 2 * i, out + 2 * i);
  return 0;
}

int ref_lorenz(double s, double r, double b, double h, uint32_t steps, const double* in, double* out, uint64_t n) {
  b12::Lorenz<3, double> m(s, r, b, h, steps);
  for (uint64_t i = 0; i < n; ++i) m(in + 3 * i, out + 3 * i);
  return 0;
}

int ref_duffing(double delta, double alpha, double beta, double gamma, double omega, double t0, double t1,
                double h, const double* in, double* out, uint64_t n) {
  b12::RK4<2, double, DuffRhs> m(DuffRhs{delta, alpha, beta, gamma, omega}, t0, t1, h);
  for (uint64_t i = 0; i < n; ++i) m(in + 2 * i, out + 2 * i);
  return 0;
}

int ref_identity(int dim, const double* in, double* out, uint64_t n) {
  if (dim != 3) return -1;
  b12::Identity<3, double> m;
  for (uint64_t i = 0; i < n; ++i) m(in + 3 * i, out + 3 * i);
  return 0;
}

}  // extern "C"
