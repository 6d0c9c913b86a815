// Device-side building blocks of the b12x hot path: box keys, domain geometry, grids, maps and
// the point -> leaf descent. Everything here is FP64 with contraction disabled at build time, and
// every expression keeps the operation order of the reference file cited beside it (paths
// relative to B12c/src/include/), so keys and targets are bit-identical to B12's host path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b12x {

constexpr int kMaxDim = 4;
constexpr uint8_t kLeafDepth = 255;      // TypeDefinitions.hpp getLeafDepth
constexpr uint8_t kUnsubDepth = 254;     // TypeDefinitions.hpp getUnsubdividableDepth
constexpr int kMaxDepth = 64;            // getMaxDepth<64>
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint64_t kInvalidKey = ~0ull;  // key of the invalid box (Box.hpp:14-19)

enum Flag : uint8_t { F_NONE = 0, F_HIT = 1, F_SD = 8 };

enum MapKind : int { MAP_IDENTITY = 0, MAP_HENON = 1, MAP_LORENZ = 2, MAP_HENON_HEILES = 3, MAP_DUFFING = 4 };
enum GridKind : int { GRID_INNER = 0, GRID_UNIFORM = 1 };

__host__ __device__ inline bool areFlagsSet(uint8_t f, uint8_t c) { return c == (f & c); }
__host__ __device__ inline bool isAnyFlagSet(uint8_t f, uint8_t c) { return (f & c) != 0; }

// BitPattern.hpp:4-28 safe shifts
__host__ __device__ inline uint64_t shl(uint64_t b, int i) { return i <= 0 ? b : (i >= 64 ? 0ull : b << i); }
__host__ __device__ inline uint64_t shr(uint64_t b, int i) { return i <= 0 ? b : (i >= 64 ? 0ull : b >> i); }

// Box.hpp:65-72 comparable key K = bits<<(64-d) | (2^(63-d) - 1). Leaves are stored by K (plus
// depth and flags); for 1 <= d <= 63 the bit pattern is K >> (64-d), for d = 64 it is K, for
// d = 0 it is 0.
__host__ __device__ inline uint64_t keyOf(uint64_t bits, int d) {
  int s = kMaxDepth - d;
  return shl(bits, s) | (shl(1ull, s - 1) - 1ull);
}
__host__ __device__ inline uint64_t bitsOf(uint64_t key, int d) { return d == 0 ? 0ull : shr(key, kMaxDepth - d); }

// ---------------------------------------------------------------------------------------------
// Domain and subdivision scheme (ImplicitBoxTree.hpp:574-587 adaptSdCount)
// ---------------------------------------------------------------------------------------------
struct Domain {
  double center[kMaxDim];
  double radius[kMaxDim];
  int isDefault;               // sdScheme[i] == i % DIM for all i
  uint8_t sdScheme[kMaxDepth + 1];
  uint8_t sdCount[(kMaxDepth + 2) * kMaxDim];  // [d*DIM + j]
};

// coordinateFunctions.hpp:7-19 isPointInBox
template <int DIM>
__device__ inline bool inDomain(const Domain& D, const double* p) {
  bool res = true;
#pragma unroll
  for (int j = 0; j < DIM; ++j) res = res && D.center[j] - D.radius[j] <= p[j] && p[j] < D.center[j] + D.radius[j];
  return res;
}

// Spread the low bits of v so that bit i lands on bit i*DIM (DIM = 1..4); bits that would land
// at or above 64 are dropped, as the position bound of coordinateFunctions.hpp:81 does.
template <int DIM>
__device__ inline uint64_t spreadBits(uint64_t v) {
  if (DIM == 1) return v;
  if (DIM == 2) {
    v &= 0xFFFFFFFFull;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
  }
  if (DIM == 3) {  // 22 bits: bits 0..20 -> 0..60, bit 21 -> 63
    const uint64_t top = (v >> 21) & 1ull;
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x001F00000000FFFFull;
    v = (v | (v << 16)) & 0x001F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v | (top << 63);
  }
  // DIM == 4: 16 bits
  v &= 0xFFFFull;
  v = (v | (v << 24)) & 0x000000FF000000FFull;
  v = (v | (v << 12)) & 0x000F000F000F000Full;
  v = (v | (v << 6)) & 0x0303030303030303ull;
  v = (v | (v << 3)) & 0x1111111111111111ull;
  return v;
}

// Gather every DIM-th bit starting at bit 0 into the low bits (inverse of spreadBits).
template <int DIM>
__device__ inline uint64_t compactBits(uint64_t v) {
  if (DIM == 1) return v;
  if (DIM == 2) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FFull;
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFFull;
    v = (v | (v >> 16)) & 0x00000000FFFFFFFFull;
    return v;
  }
  if (DIM == 3) {
    const uint64_t top = (v >> 63) & 1ull;
    v &= 0x1249249249249249ull;
    v = (v | (v >> 2)) & 0x10C30C30C30C30C3ull;
    v = (v | (v >> 4)) & 0x100F00F00F00F00Full;
    v = (v | (v >> 8)) & 0x001F0000FF0000FFull;
    v = (v | (v >> 16)) & 0x001F00000000FFFFull;
    v = (v | (v >> 32)) & 0x00000000001FFFFFull;
    return v | (top << 21);
  }
  v &= 0x1111111111111111ull;
  v = (v | (v >> 3)) & 0x0303030303030303ull;
  v = (v | (v >> 6)) & 0x000F000F000F000Full;
  v = (v | (v >> 12)) & 0x000000FF000000FFull;
  v = (v | (v >> 24)) & 0x000000000000FFFFull;
  return v;
}

// coordinateFunctions.hpp:61-93 computeBoxFromPointFixedSdScheme and :21-59 computeBoxFromPoint.
// Returns the max-depth key (= bit pattern at depth 64) or kInvalidKey outside the domain.
template <int DIM>
__device__ inline uint64_t pointKey(const Domain& D, const double* p) {
  if (!inDomain<DIM>(D, p)) return kInvalidKey;
  uint64_t res = 0;
  if (D.isDefault) {
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      const int nb = kMaxDepth / DIM + (j < kMaxDepth % DIM ? 1 : 0);
      uint64_t b = (uint64_t)floor(ldexp(((p[j] - D.center[j]) / D.radius[j] + 1.0) / 2.0, nb));
      const int shift = (kMaxDepth - 1 - j) % DIM;
      // bit i of b -> bit i*DIM + shift for i*DIM + shift < 64, i.e. i < nb
      uint64_t m = (nb >= 64) ? b : (b & ((1ull << nb) - 1ull));
      res |= spreadBits<DIM>(m) << shift;
    }
  } else {
    const uint8_t* cnt = D.sdCount + kMaxDepth * DIM;
    uint64_t bits[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j)
      bits[j] = (uint64_t)floor(ldexp(((p[j] - D.center[j]) / D.radius[j] + 1.0) / 2.0, int(cnt[j])));
    for (int i = kMaxDepth - 1; i >= 0; --i) {
      const int j = D.sdScheme[i];
#pragma unroll
      for (int jj = 0; jj < DIM; ++jj) {
        if (jj == j) {
          res |= (bits[jj] & 1ull) << (kMaxDepth - 1 - i);
          bits[jj] >>= 1;
        }
      }
    }
  }
  return res;
}

// coordinateFunctions.hpp:128-159 computeCoordinatesOfBoxFixedSdScheme and :95-126 general form:
// box (key, depth) -> centre c[] and radius rb[].
template <int DIM>
__device__ inline void boxCoords(const Domain& D, uint64_t key, int d, double* c, double* rb) {
  const uint64_t bits = bitsOf(key, d);
  if (D.isDefault) {
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      const int shift = (DIM + d - 1 - j) % DIM;
      const uint64_t bj = compactBits<DIM>(bits >> shift);
      const int nb = d / DIM + (j < d % DIM ? 1 : 0);
      rb[j] = ldexp(D.radius[j], -nb);
      c[j] = (ldexp(double(bj), -nb) * 2.0 - 1.0) * D.radius[j] + D.center[j] + rb[j];
    }
  } else {
    uint64_t bj[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) bj[j] = 0;
    for (int i = 0; i < d; ++i) {
      const int s = D.sdScheme[i];
      const uint64_t bit = (bits >> (d - 1 - i)) & 1ull;
#pragma unroll
      for (int j = 0; j < DIM; ++j)
        if (j == s) bj[j] = (bj[j] << 1) | bit;
    }
    const uint8_t* cnt = D.sdCount + d * DIM;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      rb[j] = ldexp(D.radius[j], -int(cnt[j]));
      c[j] = (ldexp(double(bj[j]), -int(cnt[j])) * 2.0 - 1.0) * D.radius[j] + D.center[j] + rb[j];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Maps. Parameters are precomputed on the host where the reference divides a member by a
// constant inside the loop (h/6, h/2, h/3): the quotient is the same IEEE value either way.
// ---------------------------------------------------------------------------------------------
struct ForcingStep {  // one step of the non-autonomous RK4 time loop (MapTools.hpp:54-56)
  double h, h6, h2, h3, f0, fh, f1;
};

struct MapParams {
  int kind;
  uint32_t steps;
  int order;
  uint32_t nForcing;
  double p[8];
  const ForcingStep* forcing;  // device pointer (Duffing)
};

// Map_Henon.h:17-22
__device__ inline void henonMap(double a, double b, const double* x, double* y) {
  y[0] = 1.0 - a * x[0] * x[0] + x[1];
  y[1] = b * x[0];
}

// Map_Lorenz.h:20-25
__device__ inline void lorenzRhs(double s, double r, double b, const double* x, double* f) {
  f[0] = s * (x[1] - x[0]);
  f[1] = r * x[0] - x[1] - x[0] * x[2];
  f[2] = -b * x[2] + x[0] * x[1];
}

// Hénon–Heiles force -grad V (build-defined; not in the reference)
__device__ inline void hhForce(const double* q, double* f) {
  f[0] = -q[0] - 2.0 * q[0] * q[1];
  f[1] = -q[1] - q[0] * q[0] + q[1] * q[1];
}

// Forced Duffing right-hand side (build-defined; not in the reference)
__device__ inline void duffingRhs(double delta, double alpha, double beta, double f, const double* x, double* r) {
  r[0] = x[1];
  r[1] = f - delta * x[1] - alpha * x[0] - beta * x[0] * x[0] * x[0];
}

template <int DIM, int MAP>
__device__ inline void applyMap(const MapParams& M, const double* in, double* out) {
  if constexpr (MAP == MAP_IDENTITY) {  // Map_Identity.h:11-20
#pragma unroll
    for (int j = 0; j < DIM; ++j) out[j] = in[j];
  } else if constexpr (MAP == MAP_HENON) {  // MapTools.hpp:4-32 IteratingMap
    const double a = M.p[0], b = M.p[1];
    uint32_t i = 0;
    if (M.steps & 1u) {
      henonMap(a, b, in, out);
      ++i;
    } else {
      out[0] = in[0];
      out[1] = in[1];
    }
    for (; i < M.steps; i += 2u) {
      double t[2];
      henonMap(a, b, out, t);
      henonMap(a, b, t, out);
    }
  } else if constexpr (MAP == MAP_LORENZ) {  // MapTools.hpp:116-182 RK4Auto
    const double s = M.p[0], r = M.p[1], b = M.p[2], h = M.p[3], h6 = M.p[4], h2 = M.p[5], h3 = M.p[6];
    double x0 = in[0], x1 = in[1], x2 = in[2];
    double r0, r1, r2;
    for (uint32_t st = 0; st < M.steps; ++st) {
      double x[3] = {x0, x1, x2}, k[3], a[3];
      lorenzRhs(s, r, b, x, k);
      r0 = x0 + h6 * k[0]; r1 = x1 + h6 * k[1]; r2 = x2 + h6 * k[2];
      a[0] = x0 + h2 * k[0]; a[1] = x1 + h2 * k[1]; a[2] = x2 + h2 * k[2];
      lorenzRhs(s, r, b, a, k);
      r0 += h3 * k[0]; r1 += h3 * k[1]; r2 += h3 * k[2];
      a[0] = x0 + h2 * k[0]; a[1] = x1 + h2 * k[1]; a[2] = x2 + h2 * k[2];
      lorenzRhs(s, r, b, a, k);
      r0 += h3 * k[0]; r1 += h3 * k[1]; r2 += h3 * k[2];
      a[0] = x0 + h * k[0]; a[1] = x1 + h * k[1]; a[2] = x2 + h * k[2];
      lorenzRhs(s, r, b, a, k);
      r0 += h6 * k[0]; r1 += h6 * k[1]; r2 += h6 * k[2];
      x0 = r0; x1 = r1; x2 = r2;
    }
    out[0] = x0; out[1] = x1; out[2] = x2;
  } else if constexpr (MAP == MAP_HENON_HEILES) {  // MapTools.hpp:369-523 SymplecticIntegratorAuto
    const double h = M.p[0], factor = M.p[1], cbrt2 = M.p[2];
    double q[2] = {in[0], in[1]}, p[2] = {in[2], in[3]}, t[2];
    for (uint32_t st = 0; st < M.steps; ++st) {
      if (M.order == 2) {
        q[0] += 0.5 * p[0] * h; q[1] += 0.5 * p[1] * h;
        hhForce(q, t);
        p[0] += t[0] * h; p[1] += t[1] * h;
        q[0] += 0.5 * p[0] * h; q[1] += 0.5 * p[1] * h;
      } else if (M.order == 1) {
        hhForce(q, t);
        p[0] += t[0] * h; p[1] += t[1] * h;
        q[0] += p[0] * h; q[1] += p[1] * h;
      } else if (M.order == 3) {
        hhForce(q, t);
        p[0] += t[0] * h; p[1] += t[1] * h;
        q[0] -= 1.0 / 24.0 * p[0] * h; q[1] -= 1.0 / 24.0 * p[1] * h;
        hhForce(q, t);
        p[0] -= 2.0 / 3.0 * t[0] * h; p[1] -= 2.0 / 3.0 * t[1] * h;
        q[0] += 3.0 / 4.0 * p[0] * h; q[1] += 3.0 / 4.0 * p[1] * h;
        hhForce(q, t);
        p[0] += 2.0 / 3.0 * t[0] * h; p[1] += 2.0 / 3.0 * t[1] * h;
        q[0] += 7.0 / 24.0 * p[0] * h; q[1] += 7.0 / 24.0 * p[1] * h;
      } else {
        hhForce(q, t);
        p[0] += 0.5 * factor * t[0] * h; p[1] += 0.5 * factor * t[1] * h;
        q[0] += factor * p[0] * h; q[1] += factor * p[1] * h;
        hhForce(q, t);
        p[0] += 0.5 * (1.0 - cbrt2) * factor * t[0] * h; p[1] += 0.5 * (1.0 - cbrt2) * factor * t[1] * h;
        q[0] -= cbrt2 * factor * p[0] * h; q[1] -= cbrt2 * factor * p[1] * h;
        hhForce(q, t);
        p[0] += 0.5 * (1.0 - cbrt2) * factor * t[0] * h; p[1] += 0.5 * (1.0 - cbrt2) * factor * t[1] * h;
        q[0] += factor * p[0] * h; q[1] += factor * p[1] * h;
        hhForce(q, t);
        p[0] += 0.5 * factor * t[0] * h; p[1] += 0.5 * factor * t[1] * h;
      }
    }
    out[0] = q[0]; out[1] = q[1]; out[2] = p[0]; out[3] = p[1];
  } else if constexpr (MAP == MAP_DUFFING) {  // MapTools.hpp:40-108 RK4 with a forcing table
    const double dl = M.p[0], al = M.p[1], be = M.p[2];
    double x[2] = {in[0], in[1]}, k[2], a[2], r[2];
    for (uint32_t st = 0; st < M.nForcing; ++st) {
      const ForcingStep S = M.forcing[st];
      duffingRhs(dl, al, be, S.f0, x, k);
      r[0] = x[0] + S.h6 * k[0]; r[1] = x[1] + S.h6 * k[1];
      a[0] = x[0] + S.h2 * k[0]; a[1] = x[1] + S.h2 * k[1];
      duffingRhs(dl, al, be, S.fh, a, k);
      r[0] += S.h3 * k[0]; r[1] += S.h3 * k[1];
      a[0] = x[0] + S.h2 * k[0]; a[1] = x[1] + S.h2 * k[1];
      duffingRhs(dl, al, be, S.fh, a, k);
      r[0] += S.h3 * k[0]; r[1] += S.h3 * k[1];
      a[0] = x[0] + S.h * k[0]; a[1] = x[1] + S.h * k[1];
      duffingRhs(dl, al, be, S.f1, a, k);
      r[0] += S.h6 * k[0]; r[1] += S.h6 * k[1];
      x[0] = r[0]; x[1] = r[1];
    }
    out[0] = x[0]; out[1] = x[1];
  }
}

// The same maps for P points at once: every integrator step runs over the P points before the
// next, so a lane carries P independent FP64 chains. Each point's operations and their order are
// those of applyMap.
template <int DIM, int MAP, int P>
__device__ inline void applyMapN(const MapParams& M, double (&x)[P][DIM], double (&y)[P][DIM]) {
  if constexpr (MAP == MAP_LORENZ) {  // MapTools.hpp:116-182 RK4Auto
    const double s = M.p[0], r = M.p[1], b = M.p[2], h = M.p[3], h6 = M.p[4], h2 = M.p[5], h3 = M.p[6];
    double z[P][3], rr[P][3], a[P][3], k[P][3];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int j = 0; j < 3; ++j) z[q][j] = x[q][j];
    for (uint32_t st = 0; st < M.steps; ++st) {
#pragma unroll
      for (int q = 0; q < P; ++q) lorenzRhs(s, r, b, z[q], k[q]);
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          rr[q][j] = z[q][j] + h6 * k[q][j];
          a[q][j] = z[q][j] + h2 * k[q][j];
        }
#pragma unroll
      for (int q = 0; q < P; ++q) lorenzRhs(s, r, b, a[q], k[q]);
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          rr[q][j] += h3 * k[q][j];
          a[q][j] = z[q][j] + h2 * k[q][j];
        }
#pragma unroll
      for (int q = 0; q < P; ++q) lorenzRhs(s, r, b, a[q], k[q]);
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          rr[q][j] += h3 * k[q][j];
          a[q][j] = z[q][j] + h * k[q][j];
        }
#pragma unroll
      for (int q = 0; q < P; ++q) lorenzRhs(s, r, b, a[q], k[q]);
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j) z[q][j] = rr[q][j] + h6 * k[q][j];
    }
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int j = 0; j < 3; ++j) y[q][j] = z[q][j];
  } else if constexpr (MAP == MAP_HENON_HEILES) {
    if (M.order != 2) {  // orders 1, 3, 4 point by point
#pragma unroll
      for (int q = 0; q < P; ++q) applyMap<DIM, MAP>(M, x[q], y[q]);
      return;
    }
    const double h = M.p[0];  // MapTools.hpp:399-420 Störmer-Verlet
    double qq[P][2], pp[P][2], t[P][2];
#pragma unroll
    for (int q = 0; q < P; ++q) {
      qq[q][0] = x[q][0]; qq[q][1] = x[q][1]; pp[q][0] = x[q][2]; pp[q][1] = x[q][3];
    }
    for (uint32_t st = 0; st < M.steps; ++st) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        qq[q][0] += 0.5 * pp[q][0] * h; qq[q][1] += 0.5 * pp[q][1] * h;
      }
#pragma unroll
      for (int q = 0; q < P; ++q) hhForce(qq[q], t[q]);
#pragma unroll
      for (int q = 0; q < P; ++q) {
        pp[q][0] += t[q][0] * h; pp[q][1] += t[q][1] * h;
        qq[q][0] += 0.5 * pp[q][0] * h; qq[q][1] += 0.5 * pp[q][1] * h;
      }
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
      y[q][0] = qq[q][0]; y[q][1] = qq[q][1]; y[q][2] = pp[q][0]; y[q][3] = pp[q][1];
    }
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) applyMap<DIM, MAP>(M, x[q], y[q]);
  }
}

// ---------------------------------------------------------------------------------------------
// Grids (Grids.hpp:99-176). InnerGrid offsets depend only on i mod k, so they come from a table
// the host fills with the reference's expression. UniformGrid follows the host branch:
// std::mt19937(seed + i) and uniform_real_distribution<double>(-1, 1), evaluated lazily (only
// the 2*DIM state words that the first 2*DIM outputs need are twisted).
// ---------------------------------------------------------------------------------------------
struct GridParams {
  int kind;
  uint32_t k;           // points per box
  uint64_t seed;
  const double* table;  // k*DIM offsets (InnerGrid)
};

template <int DIM>
__device__ inline void uniformGridPoint(uint64_t seed, uint64_t i, double* g) {
  constexpr int NOUT = 2 * DIM;
  uint32_t lo[NOUT + 1], hi[NOUT];
  uint32_t x = uint32_t(seed + i);
  lo[0] = x;
#pragma unroll
  for (int t = 1; t <= NOUT; ++t) {
    x = 1812433253u * (x ^ (x >> 30)) + uint32_t(t);
    lo[t] = x;
  }
  for (int t = NOUT + 1; t < 397; ++t) x = 1812433253u * (x ^ (x >> 30)) + uint32_t(t);
#pragma unroll
  for (int t = 0; t < NOUT; ++t) {
    x = 1812433253u * (x ^ (x >> 30)) + uint32_t(397 + t);
    hi[t] = x;
  }
#pragma unroll
  for (int j = 0; j < DIM; ++j) {
    double sum = 0.0, tmp = 1.0;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = 2 * j + u;
      uint32_t y = (lo[t] & 0x80000000u) | (lo[t + 1] & 0x7fffffffu);
      y = hi[t] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      sum += double(y) * tmp;
      tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;  // nextafter(1, 0)
    g[j] = (ret * 2.0) + -1.0;
  }
}

// ---------------------------------------------------------------------------------------------
// Descent: lower_bound of the point key over the sorted leaf keys (ImplicitBoxTree.hpp:2397)
// followed by ComputeSearchIndexFunctor (Functors.hpp:287-315) at leaf depth. A valid point has
// depth 64, so isStrictlyPreceding(leaf, point) reduces to K(leaf) < K(point).
// ---------------------------------------------------------------------------------------------
__device__ inline bool leafContains(uint64_t pk, uint64_t lk, int ld) {
  return ld == 0 || ((pk ^ lk) >> (kMaxDepth - ld)) == 0ull;
}

__device__ inline uint32_t locate(const uint64_t* __restrict__ keys, const uint8_t* __restrict__ depth, uint32_t n,
                                  uint64_t pk) {
  if (pk == kInvalidKey) {
    // invalid points sort past every leaf; hasSamePath(invalid, leaf) holds iff the leaf's bit
    // pattern is zero (Box.hpp:74-83), which B12 then reports as a hit on leaf n-1
    const uint32_t l = n - 1;
    return bitsOf(keys[l], depth[l]) == 0ull ? l : kInvalid;
  }
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (keys[lo + half] < pk) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const uint32_t s = lo;
  const uint32_t i1 = s > 0 ? s - 1 : 0;
  const uint32_t i2 = s < n - 1 ? s : n - 1;
  if (leafContains(pk, keys[i1], depth[i1])) return i1;
  if (leafContains(pk, keys[i2], depth[i2])) return s;
  return kInvalid;
}

// ---------------------------------------------------------------------------------------------
// Boxes at a depth d between the root and the leaves, or at the unsubdividable depth
// (ImplicitBoxTree.hpp:1383-1473, 1990-2021, Box.hpp:159-174, 200-204): a leaf truncated to d
// (decrementDepthTo; at the unsubdividable depth its parent) and the search of a point among the
// boxes at d (ImplicitBoxTree.hpp:2433-2566: ComputeSearchIndexFunctor on the two neighbours of
// the lower bound, truncated to d and admitted by their `in` flag, then the leaf's box index).
// ---------------------------------------------------------------------------------------------
__device__ inline void truncateTo(uint64_t key, int depth, uint8_t d, uint64_t& bits, int& td) {
  bits = bitsOf(key, depth);
  td = depth;
  if (d < depth) {
    bits >>= (depth - d);
    td = d;
  } else if (d == kUnsubDepth) {
    bits >>= 1;
    td = depth > 0 ? depth - 1 : 0;
  }
}

// hasSamePath(point, box) for a point key (an invalid point matches an all-zero pattern only)
__device__ inline bool pointInBox(uint64_t pk, uint64_t bits, int td) {
  if (pk == kInvalidKey) return bits == 0ull;
  return td == 0 || ((pk ^ keyOf(bits, td)) >> (kMaxDepth - td)) == 0ull;
}

__device__ inline uint32_t locateInDepth(const uint64_t* __restrict__ keys, const uint8_t* __restrict__ depth,
                                         uint32_t n, uint64_t pk, uint8_t d, const uint8_t* __restrict__ in,
                                         const uint32_t* __restrict__ box) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (keys[lo + half] < pk) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const uint32_t s = lo;
  const uint32_t i1 = s > 0 ? s - 1 : 0;
  const uint32_t i2 = s < n - 1 ? s : n - 1;
  uint64_t b;
  int td;
  if (in[i1]) {
    truncateTo(keys[i1], depth[i1], d, b, td);
    if (pointInBox(pk, b, td)) return box[i1];
  }
  if (in[i2]) {
    truncateTo(keys[i2], depth[i2], d, b, td);
    if (pointInBox(pk, b, td)) return box[s < n - 1 ? s : n - 1];
  }
  return kInvalid;
}

// locate() for P keys at once: a lower bound whose trip count depends on n only (so the P
// searches run in lockstep, each probe of every key in flight together), then locate()'s leaf
// checks and its invalid-point case. n >= 1.
template <int P>
__device__ inline void locateN(const uint64_t* __restrict__ keys, const uint8_t* __restrict__ depth, uint32_t n,
                               const uint64_t (&pk)[P], uint32_t (&out)[P]) {
  uint32_t base[P];
#pragma unroll
  for (int q = 0; q < P; ++q) base[q] = 0;
  uint32_t len = n;
  while (len > 1) {
    const uint32_t half = len >> 1;
    uint64_t kv[P];
#pragma unroll
    for (int q = 0; q < P; ++q) kv[q] = keys[base[q] + half];
#pragma unroll
    for (int q = 0; q < P; ++q) base[q] = kv[q] < pk[q] ? base[q] + half : base[q];
    len -= half;
  }
  uint64_t kb[P];
#pragma unroll
  for (int q = 0; q < P; ++q) kb[q] = keys[base[q]];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const uint32_t s = base[q] + (kb[q] < pk[q] ? 1u : 0u);  // lower bound
    if (pk[q] == kInvalidKey) {
      const uint32_t l = n - 1;
      out[q] = bitsOf(keys[l], depth[l]) == 0ull ? l : kInvalid;
      continue;
    }
    const uint32_t i1 = s > 0 ? s - 1 : 0;
    const uint32_t i2 = s < n - 1 ? s : n - 1;
    if (leafContains(pk[q], keys[i1], depth[i1])) out[q] = i1;
    else if (leafContains(pk[q], keys[i2], depth[i2])) out[q] = s;
    else out[q] = kInvalid;
  }
}

// Energy of the Hénon-Heiles Hamiltonian over a box (build-defined band predicate; no reference
// counterpart): H(q, p) = (p1^2 + p2^2)/2 + (q1^2 + q2^2)/2 + q1^2 q2 - q2^3/3 on
// [c - r, c + r] in (q1, q2, p1, p2), as an interval in round-to-nearest FP64: squares by sign
// case, q1^2 q2 as the hull of the four corner products, q2^3 by monotonicity, summed in a
// fixed order.
__device__ inline void intervalSquare(double a, double b, double& lo, double& hi) {
  if (a >= 0.0) {
    lo = a * a;
    hi = b * b;
  } else if (b <= 0.0) {
    lo = b * b;
    hi = a * a;
  } else {
    const double x = a * a, y = b * b;
    lo = 0.0;
    hi = x > y ? x : y;
  }
}

__device__ inline bool henonHeilesBand(const double* c, const double* r, double emin, double emax) {
  const double q1a = c[0] - r[0], q1b = c[0] + r[0], q2a = c[1] - r[1], q2b = c[1] + r[1];
  const double p1a = c[2] - r[2], p1b = c[2] + r[2], p2a = c[3] - r[3], p2b = c[3] + r[3];
  double Q1lo, Q1hi, Q2lo, Q2hi, P1lo, P1hi, P2lo, P2hi;
  intervalSquare(q1a, q1b, Q1lo, Q1hi);
  intervalSquare(q2a, q2b, Q2lo, Q2hi);
  intervalSquare(p1a, p1b, P1lo, P1hi);
  intervalSquare(p2a, p2b, P2lo, P2hi);
  const double m1 = Q1lo * q2a, m2 = Q1lo * q2b, m3 = Q1hi * q2a, m4 = Q1hi * q2b;
  const double lo12 = m1 < m2 ? m1 : m2, lo34 = m3 < m4 ? m3 : m4;
  const double hi12 = m1 > m2 ? m1 : m2, hi34 = m3 > m4 ? m3 : m4;
  const double Mlo = lo12 < lo34 ? lo12 : lo34, Mhi = hi12 > hi34 ? hi12 : hi34;
  const double c3lo = (q2a * q2a) * q2a, c3hi = (q2b * q2b) * q2b;
  const double lo = ((0.5 * (P1lo + P2lo) + 0.5 * (Q1lo + Q2lo)) + Mlo) - c3hi / 3.0;
  const double hi = ((0.5 * (P1hi + P2hi) + 0.5 * (Q1hi + Q2hi)) + Mhi) - c3lo / 3.0;
  return hi >= emin && lo <= emax;
}

// Descent with the upper levels staged in LDS. sm[0..S) holds the leaf keys at positions
// pos(j) = floor(j * n / S), S = min(kDescentSamples, n) (sampleKeysKernel writes them). The
// count j of samples below the point key brackets the lower bound: it is 0 when j = 0, else it
// lies in (pos(j-1), pos(j)] (pos(S) = n), where the global search finishes. S = n makes the LDS
// search exact. Same result as locate().
constexpr uint32_t kDescentSamples = 2048;

__device__ inline uint32_t samplePos(uint32_t j, uint32_t n, uint32_t S) {
  return uint32_t((uint64_t(j) * n) / S);
}

__device__ inline uint32_t locateStaged(const uint64_t* __restrict__ keys, const uint8_t* __restrict__ depth,
                                        uint32_t n, uint64_t pk, const uint64_t* sm, uint32_t S) {
  if (pk == kInvalidKey) {
    const uint32_t l = n - 1;
    return bitsOf(keys[l], depth[l]) == 0ull ? l : kInvalid;
  }
  uint32_t lo = 0, len = S;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (sm[lo + half] < pk) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  uint32_t s = 0;
  if (lo > 0) {
    s = samplePos(lo - 1, n, S) + 1;
    len = (lo < S ? samplePos(lo, n, S) : n) - s;
    while (len > 0) {
      const uint32_t half = len >> 1;
      if (keys[s + half] < pk) {
        s += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
  }
  const uint32_t i1 = s > 0 ? s - 1 : 0;
  const uint32_t i2 = s < n - 1 ? s : n - 1;
  if (leafContains(pk, keys[i1], depth[i1])) return i1;
  if (leafContains(pk, keys[i2], depth[i2])) return s;
  return kInvalid;
}

}  // namespace b12x
