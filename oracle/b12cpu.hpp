// b12cpu — CPU restatement of B12's set-oriented hot path.
//
// TEST INFRASTRUCTURE ONLY. This header is the oracle that the HIP product is
// checked against. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it (through liboracle_b12cpu.so); the product
// library never includes, links or calls it.
//
// It restates, in plain C++17 with FP contraction off, the host (Thrust CPP/OMP)
// semantics of the reference files cited per function below (paths relative to
// B12c/src/include/). Every FP64 expression keeps the reference's operation
// order so that covers, leaf keys and hit counts are bit-exact.
//
// Parity pins (see tests/): integrators are pinned bit-for-bit against the
// reference's own Map headers compiled into oracle/_ref; key algebra, descent and
// tree operations against closed forms and brute force on small trees; UniformGrid
// against libstdc++'s std::mt19937 / uniform_real_distribution; power iteration
// against numpy/scipy eigensolvers. Functions with no such check say
// "parity unpinned" in their comment.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace b12cpu {

using Depth = uint8_t;      // TypeDefinitions.h:19
using Flags = uint8_t;      // Flags.h:9
using NrBoxes = uint32_t;   // TypeDefinitions.h:45
using NrPoints = uint64_t;  // TypeDefinitions.h:57

constexpr Depth kLeafDepth = 255;        // TypeDefinitions.hpp:4-7 getLeafDepth
constexpr Depth kUnsubDepth = 254;       // TypeDefinitions.hpp:9-12 getUnsubdividableDepth
constexpr Depth kMaxDepth = 64;          // TypeDefinitions.hpp:14-19 getMaxDepth<64>
constexpr NrBoxes kInvalid = 0xFFFFFFFFu; // TypeDefinitions.hpp:21-26

// Flags.h:11-22
enum : Flags { NONE = 0x00, HIT = 0x01, INS = 0x02, EXPD = 0x04, SD = 0x08,
               EQU = 0x10, ENTR = 0x20, EXTR = 0x40, ALL = 0xFF };

// Flags.hpp:9-37
inline bool areFlagsSet(Flags f, Flags c) { return c == (f & c); }
inline bool isAnyFlagSet(Flags f, Flags c) { return (f & c) != 0; }
inline Flags setF(Flags f, Flags s) { return f | s; }
inline Flags unsetF(Flags f, Flags u) { return f & Flags(~u); }
inline Flags changeF(Flags f, Flags u, Flags s) { return areFlagsSet(f, u) ? setF(unsetF(f, u), s) : f; }

struct Box {
  uint64_t bits;
  Depth depth;
  Flags flags;
};

// BitPattern.hpp:4-28
inline uint64_t shl(uint64_t b, int i) { return i <= 0 ? b : (i >= 64 ? 0ull : b << i); }
inline uint64_t shr(uint64_t b, int i) { return i <= 0 ? b : (i >= 64 ? 0ull : b >> i); }

// Box.hpp:7-19, 74-83
inline Box surroundingBox() { return {0ull, 0, NONE}; }
inline Box invalidBox() { return {~0ull, kLeafDepth, NONE}; }
inline bool isInvalid(const Box& b) { return b.depth > kMaxDepth; }
inline bool isSurrounding(const Box& b) { return b.depth == 0 && b.bits == 0; }

// Box.hpp:65-72 getComparableBitPattern
inline uint64_t comparableKey(const Box& b) {
  int d = int(kMaxDepth) - int(b.depth);
  return shl(b.bits, d) | (shl(1ull, d - 1) - 1ull);
}

// Box.hpp:74-83 hasSamePath
inline bool hasSamePath(const Box& a, const Box& b) {
  int diff = int(a.depth) - int(b.depth);
  return (isInvalid(a) && isInvalid(b)) || shr(a.bits, diff) == shr(b.bits, -diff);
}

// Box.hpp:113-120 isSameBox
inline bool isSameBox(const Box& a, const Box& b) {
  return (isInvalid(a) && isInvalid(b)) || (a.bits == b.bits && a.depth == b.depth);
}

// Box.hpp:122-135 isStrictlyPreceding
inline bool isStrictlyPreceding(const Box& a, const Box& b) {
  uint64_t ka = comparableKey(a), kb = comparableKey(b);
  return !isInvalid(a) && (isInvalid(b) || ka < kb || (ka == kb && b.depth < a.depth));
}

// Box.hpp:206-220 computeDepthOfMutualAncestor: leading equal bits of the two boxes' patterns
// at maximal depth, at most the smaller depth
inline Depth mutualAncestorDepth(const Box& a, const Box& b) {
  const uint64_t x = shl(a.bits, int(kMaxDepth) - int(a.depth)) ^ shl(b.bits, int(kMaxDepth) - int(b.depth));
  const int lz = x ? __builtin_clzll(x) : 64;
  return Depth(std::min<int>(lz, std::min(a.depth, b.depth)));
}

// Box.hpp:178-198 changeDepthTo (leaf depth: unchanged; deeper: padded with zero bits)
inline Box changeDepthTo(const Box& b, Depth d) {
  Box r = b;
  if (d == kUnsubDepth) {
    r.bits = shr(b.bits, 1);
    r.depth = Depth(b.depth - (b.depth > 0 ? 1 : 0));
  } else if (d < b.depth) {
    r.bits = shr(b.bits, int(b.depth) - int(d));
    r.depth = d;
  } else if (d > b.depth && d != kLeafDepth) {
    d = std::min<Depth>(d, kMaxDepth);
    r.bits = shl(b.bits, int(d) - int(b.depth));
    r.depth = d;
  }
  return r;
}

// Box.hpp:138-147 incrementDepthBy
inline Box incrementDepthBy(const Box& b, Depth d) {
  d = std::min<Depth>(d, Depth(kMaxDepth - b.depth));
  return {shl(b.bits, d), Depth(b.depth + d), b.flags};
}

// Box.hpp:149-157 decrementDepthBy
inline Box decrementDepthBy(const Box& b, Depth d) {
  d = std::min<Depth>(d, b.depth);
  return {shr(b.bits, d), Depth(b.depth - d), b.flags};
}

// Box.hpp:159-174 decrementDepthTo
inline Box decrementDepthTo(const Box& b, Depth d) {
  Box r = b;
  if (d < b.depth) {
    r.bits = shr(b.bits, int(b.depth) - int(d));
    r.depth = d;
  } else if (d == kUnsubDepth) {
    r.bits = shr(b.bits, 1);
    r.depth = Depth(r.depth - (b.depth > 0 ? 1 : 0));
  }
  return r;
}

// Box.hpp:200-204 hasAtLeastDepth
inline bool hasAtLeastDepth(const Box& b, Depth d) {
  return b.depth >= d || d == kUnsubDepth || d == kLeafDepth;
}

// Box.hpp:225-239
inline bool isRightHalf(const Box& b) { return (b.bits & 1ull) != 0; }
inline Box computeOtherHalf(const Box& b) {
  return {b.bits ^ uint64_t(!isSurrounding(b) && !isInvalid(b)), b.depth, b.flags};
}

// ---------------------------------------------------------------------------------------------
// Domain geometry: centre, radius, subdivision scheme (ImplicitBoxTree.hpp:27-45, 574-587,
// 1972-1979, 3512-3552)
// ---------------------------------------------------------------------------------------------
template <int DIM>
struct Geometry {
  double center[DIM];
  double radius[DIM];
  std::vector<uint8_t> sdScheme;   // sdScheme[i] = coordinate bisected at level i
  std::vector<uint8_t> sdCount;    // sdCount[d*DIM + j] = bisections of j above depth d
  bool isDefault = true;

  Geometry() {
    for (int j = 0; j < DIM; ++j) { center[j] = 0.5; radius[j] = 0.5; }
    sdScheme.resize(kMaxDepth);
    for (int i = 0; i < kMaxDepth; ++i) sdScheme[i] = uint8_t(i % DIM);
    adapt();
  }

  // ImplicitBoxTree.hpp:3512-3552 setSdScheme (periodic extension to maxDepth+1 entries)
  bool setScheme(const uint8_t* s, int n) {
    for (int i = 0; i < n; ++i) if (s[i] >= DIM) return false;
    sdScheme.assign(kMaxDepth + 1, 0);
    if (n > 0) {
      for (int i = 0; i < kMaxDepth + 1; ++i) sdScheme[i] = s[i % n];
    } else {
      for (int i = 0; i < kMaxDepth + 1; ++i) sdScheme[i] = uint8_t(i % DIM);
    }
    adapt();
    return true;
  }

  // ImplicitBoxTree.hpp:574-587 adaptSdCount; :1972-1979 isSdSchemeDefault
  void adapt() {
    sdCount.assign((kMaxDepth + 2) * DIM, 0);
    for (size_t i = 0; i < sdScheme.size(); ++i) {
      for (int j = 0; j < DIM; ++j) sdCount[(i + 1) * DIM + j] = sdCount[i * DIM + j];
      ++sdCount[(i + 1) * DIM + sdScheme[i]];
    }
    isDefault = true;
    for (size_t i = 0; i < sdScheme.size(); ++i)
      if (sdScheme[i] != uint8_t(uint8_t(i) % DIM)) isDefault = false;
  }

  // coordinateFunctions.hpp:7-19 isPointInBox
  bool inDomain(const double* p) const {
    bool res = true;
    for (int j = 0; j < DIM; ++j) res = res && center[j] - radius[j] <= p[j] && p[j] < center[j] + radius[j];
    return res;
  }

  // coordinateFunctions.hpp:61-93 computeBoxFromPointFixedSdScheme
  Box boxFromPointFixed(const double* p, Flags flags) const {
    Box res{0ull, kMaxDepth, flags};
    if (!inDomain(p)) return invalidBox();
    for (int j = 0; j < DIM; ++j) {
      int nb = kMaxDepth / DIM + (j < kMaxDepth % DIM ? 1 : 0);
      uint64_t b = (uint64_t)std::floor(std::ldexp(((p[j] - center[j]) / radius[j] + 1.0) / 2.0, nb));
      int initialShift = (kMaxDepth - 1 - j) % DIM;
      for (int i = 0; i * DIM + initialShift < kMaxDepth; ++i)
        res.bits |= shl(b, i * (DIM - 1) + initialShift) & shl(1ull, i * DIM + initialShift);
    }
    return res;
  }

  // coordinateFunctions.hpp:21-59 computeBoxFromPoint (general subdivision scheme)
  Box boxFromPointGeneral(const double* p, Flags flags) const {
    Box res{0ull, kMaxDepth, flags};
    if (!inDomain(p)) return invalidBox();
    const uint8_t* cnt = sdCount.data() + kMaxDepth * DIM;
    uint64_t bits[DIM];
    for (int j = 0; j < DIM; ++j)
      bits[j] = (uint64_t)std::floor(std::ldexp(((p[j] - center[j]) / radius[j] + 1.0) / 2.0, cnt[j]));
    for (int i = kMaxDepth - 1; i >= 0; --i) {
      for (int j = 0; j < DIM; ++j) {
        bool t = sdScheme[i] == j;
        res.bits |= shl(uint64_t(t) & bits[j], kMaxDepth - 1 - i);
        bits[j] = shr(bits[j], int(t));
      }
    }
    return res;
  }

  Box boxFromPoint(const double* p, Flags flags) const {
    return isDefault ? boxFromPointFixed(p, flags) : boxFromPointGeneral(p, flags);
  }

  // coordinateFunctions.hpp:128-159 computeCoordinatesOfBoxFixedSdScheme -> centre c, radius rb
  void coordsFixed(const Box& box, double* c, double* rb) const {
    int d = box.depth;
    for (int j = 0; j < DIM; ++j) {
      uint64_t bits = 0;
      int initialShift = (DIM + d - 1 - j) % DIM;
      for (int i = 0; i * DIM < kMaxDepth; ++i)
        bits |= shr(box.bits, i * (DIM - 1) + initialShift) & shl(1ull, i);
      int nb = d / DIM + (j < d % DIM ? 1 : 0);
      rb[j] = std::ldexp(radius[j], -nb);
      c[j] = (std::ldexp(double(bits), -nb) * 2.0 - 1.0) * radius[j] + center[j] + rb[j];
    }
  }

  // coordinateFunctions.hpp:95-126 computeCoordinatesOfBox (general scheme)
  void coordsGeneral(const Box& box, double* c, double* rb) const {
    uint64_t bits[DIM] = {};
    int d = box.depth;
    const uint8_t* cnt = sdCount.data() + d * DIM;
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < DIM; ++j) {
        bool t = sdScheme[i] == j;
        bits[j] = shl(bits[j], int(t)) | (uint64_t(t) & shr(box.bits, d - 1 - i));
      }
    for (int j = 0; j < DIM; ++j) {
      rb[j] = std::ldexp(radius[j], -int(cnt[j]));
      c[j] = (std::ldexp(double(bits[j]), -int(cnt[j])) * 2.0 - 1.0) * radius[j] + center[j] + rb[j];
    }
  }

  void coords(const Box& box, double* c, double* rb) const {
    if (isDefault) coordsFixed(box, c, rb); else coordsGeneral(box, c, rb);
  }
};

// ---------------------------------------------------------------------------------------------
// Random numbers for UniformGrid: std::mt19937 + uniform_real_distribution<double>(-1,1) as in
// libstdc++ (bits/random.tcc generate_canonical; bits/random.h:1870), restated so the result
// does not depend on the host's library. Pinned against std:: in tests/.
// ---------------------------------------------------------------------------------------------
struct MT19937 {
  uint32_t x[624];
  int idx;
  explicit MT19937(uint64_t seed) {
    x[0] = uint32_t(seed);
    for (int i = 1; i < 624; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + uint32_t(i);
    idx = 624;
  }
  void twist() {
    for (int k = 0; k < 624; ++k) {
      uint32_t y = (x[k] & 0x80000000u) | (x[(k + 1) % 624] & 0x7fffffffu);
      x[k] = x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    idx = 0;
  }
  uint32_t operator()() {
    if (idx >= 624) twist();
    uint32_t y = x[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
};

// generate_canonical<double,53>(mt19937) then *(b-a)+a with a=-1, b=1
inline double uniformMinus1To1(MT19937& g) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += double(g()) * tmp;
    tmp *= 4294967296.0;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
  return (ret * 2.0) + -1.0;
}

// ---------------------------------------------------------------------------------------------
// Grids (Grids.hpp:99-176). kind 0 = InnerGrid(n), kind 1 = UniformGrid(k, seed) host semantics.
// ---------------------------------------------------------------------------------------------
enum GridKind : int { GRID_INNER = 0, GRID_UNIFORM = 1 };

template <int DIM>
struct Grid {
  int kind = GRID_INNER;
  uint64_t n = 1;      // InnerGrid points per dimension
  uint64_t k = 1;      // points per box
  uint64_t seed = 0;   // UniformGrid seed

  uint64_t pointsPerBox() const { return k; }

  // Grids.hpp:112-117 InnerGrid::operator(); Grids.hpp:152-158 UniformGrid host branch
  void operator()(uint64_t i, double* out) const {
    if (kind == GRID_INNER) {
      uint64_t npd = 1;
      for (int j = 0; j < DIM; ++j) {
        out[j] = (double((i / npd) % n) * 2.0 + 1.0 - double(n)) / double(n);
        npd *= n;
      }
    } else {
      MT19937 g(seed + i);
      for (int j = 0; j < DIM; ++j) out[j] = uniformMinus1To1(g);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// Maps (MapTools.hpp, Map_Henon.h, Map_Lorenz.h, Map_Identity.h; the Hénon–Heiles and forced
// Duffing right-hand sides are defined by this build: they have no counterpart in the reference)
// ---------------------------------------------------------------------------------------------
enum MapKind : int { MAP_IDENTITY = 0, MAP_HENON = 1, MAP_LORENZ = 2, MAP_HENON_HEILES = 3,
                     MAP_DUFFING = 4 };

// Forcing table for the non-autonomous RK4 (MapTools.hpp:54-105): the time loop
// for (t = t0; t < tEnd; t += h_) { h = min(h_, tEnd - t); ... t, t + h/2, t + h ... }
// is state independent, so it is evaluated once; f = gamma*cos(omega*time).
struct ForcingStep {
  double h, h6, h2, h3, f0, fh, f1;
};

inline std::vector<ForcingStep> buildForcingTable(double t0, double tEnd, double hIn, double gamma,
                                                  double omega) {
  std::vector<ForcingStep> tab;
  for (double t = t0; t < tEnd; t += hIn) {
    double h = (tEnd - t) < hIn ? (tEnd - t) : hIn;
    ForcingStep s;
    s.h = h;
    s.h6 = h / 6.0;
    s.h2 = h / 2.0;
    s.h3 = h / 3.0;
    s.f0 = gamma * std::cos(omega * t);
    s.fh = gamma * std::cos(omega * (t + h / 2.0));
    s.f1 = gamma * std::cos(omega * (t + h));
    tab.push_back(s);
    if (tab.size() > 100000000) break;
  }
  return tab;
}

template <int DIM>
struct Map {
  int kind = MAP_IDENTITY;
  // Henon: a, b; Lorenz: sigma, rho, beta; HH: -; Duffing: delta, alpha, beta
  double p0 = 0, p1 = 0, p2 = 0;
  double h = 0;
  uint32_t steps = 1;
  int order = 2;  // symplectic order
  double cbrt2 = 0, factor = 0;  // order-4 constants (host cbrt)
  std::vector<ForcingStep> forcing;

  // Map_Henon.h:17-22 HenonMap::operator()
  void henon(const double* x, double* y) const {
    y[0] = 1.0 - p0 * x[0] * x[0] + x[1];
    y[1] = p1 * x[0];
  }

  // Map_Lorenz.h:20-25 LorenzRHS::operator()
  void lorenzRhs(const double* x, double* f) const {
    f[0] = p0 * (x[1] - x[0]);
    f[1] = p1 * x[0] - x[1] - x[0] * x[2];
    f[2] = -p2 * x[2] + x[0] * x[1];
  }

  // Hénon–Heiles V(q) = (q1^2+q2^2)/2 + q1^2 q2 - q2^3/3: F2(q) = -grad V (build-defined)
  static void hhForce(const double* q, double* f) {
    f[0] = -q[0] - 2.0 * q[0] * q[1];
    f[1] = -q[1] - q[0] * q[0] + q[1] * q[1];
  }

  // Forced Duffing x'' + delta x' + alpha x + beta x^3 = gamma cos(omega t) (build-defined)
  void duffingRhs(double f, const double* x, double* r) const {
    r[0] = x[1];
    r[1] = f - p0 * x[1] - p1 * x[0] - p2 * x[0] * x[0] * x[0];
  }

  void operator()(const double* in, double* out) const {
    switch (kind) {
      case MAP_IDENTITY:  // Map_Identity.h:11-20
        for (int j = 0; j < DIM; ++j) out[j] = in[j];
        break;
      case MAP_HENON: {  // MapTools.hpp:4-32 IteratingMap
        uint32_t i = 0;
        if (steps & 1u) {
          henon(in, out);
          ++i;
        } else {
          for (int j = 0; j < DIM; ++j) out[j] = in[j];
        }
        for (; i < steps; i += 2u) {
          double t[2];
          henon(out, t);
          henon(t, out);
        }
        break;
      }
      case MAP_LORENZ: {  // MapTools.hpp:116-182 RK4Auto::operator()
        double x[3], k[3], a[3];
        for (int j = 0; j < 3; ++j) x[j] = in[j];
        for (uint32_t s = 0; s < steps; ++s) {
          lorenzRhs(x, k);
          for (int j = 0; j < 3; ++j) out[j] = x[j] + h / 6.0 * k[j];
          for (int j = 0; j < 3; ++j) a[j] = x[j] + h / 2.0 * k[j];
          lorenzRhs(a, k);
          for (int j = 0; j < 3; ++j) out[j] += h / 3.0 * k[j];
          for (int j = 0; j < 3; ++j) a[j] = x[j] + h / 2.0 * k[j];
          lorenzRhs(a, k);
          for (int j = 0; j < 3; ++j) out[j] += h / 3.0 * k[j];
          for (int j = 0; j < 3; ++j) a[j] = x[j] + h * k[j];
          lorenzRhs(a, k);
          for (int j = 0; j < 3; ++j) out[j] += h / 6.0 * k[j];
          for (int j = 0; j < 3; ++j) x[j] = out[j];
        }
        break;
      }
      case MAP_HENON_HEILES: {  // MapTools.hpp:369-523 SymplecticIntegratorAuto<4>
        double* q = out;
        double* p = out + 2;
        double t[2];
        for (int j = 0; j < 4; ++j) out[j] = in[j];
        for (uint32_t s = 0; s < steps; ++s) {
          if (order == 1) {
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += t[j] * h;
          } else if (order == 2) {
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += 0.5 * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += 0.5 * t[j] * h;
          } else if (order == 3) {
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] -= 1.0 / 24.0 * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] -= 2.0 / 3.0 * t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += 3.0 / 4.0 * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += 2.0 / 3.0 * t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += 7.0 / 24.0 * t[j] * h;
          } else {
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += 0.5 * factor * t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += factor * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += 0.5 * (1.0 - cbrt2) * factor * t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] -= cbrt2 * factor * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += 0.5 * (1.0 - cbrt2) * factor * t[j] * h;
            for (int j = 0; j < 2; ++j) t[j] = p[j];
            for (int j = 0; j < 2; ++j) q[j] += factor * t[j] * h;
            hhForce(q, t);
            for (int j = 0; j < 2; ++j) p[j] += 0.5 * factor * t[j] * h;
          }
        }
        break;
      }
      case MAP_DUFFING: {  // MapTools.hpp:40-108 RK4::operator() with the forcing table
        double x[2], k[2], a[2];
        for (int j = 0; j < 2; ++j) x[j] = in[j];
        for (const ForcingStep& st : forcing) {
          duffingRhs(st.f0, x, k);
          for (int j = 0; j < 2; ++j) out[j] = x[j] + st.h6 * k[j];
          for (int j = 0; j < 2; ++j) a[j] = x[j] + st.h2 * k[j];
          duffingRhs(st.fh, a, k);
          for (int j = 0; j < 2; ++j) out[j] += st.h3 * k[j];
          for (int j = 0; j < 2; ++j) a[j] = x[j] + st.h2 * k[j];
          duffingRhs(st.fh, a, k);
          for (int j = 0; j < 2; ++j) out[j] += st.h3 * k[j];
          for (int j = 0; j < 2; ++j) a[j] = x[j] + st.h * k[j];
          duffingRhs(st.f1, a, k);
          for (int j = 0; j < 2; ++j) out[j] += st.h6 * k[j];
          for (int j = 0; j < 2; ++j) x[j] = out[j];
        }
        break;
      }
      default:
        throw std::runtime_error("unknown map kind");
    }
  }
};

// ---------------------------------------------------------------------------------------------
// COO matrix (CooMatrix.h:56-121)
// ---------------------------------------------------------------------------------------------
struct Coo {
  std::vector<NrBoxes> rows, cols;
  std::vector<double> vals;
  NrBoxes nRows = 0, nCols = 0;
};

// CooMatrix.hpp:14-56 compressIndexVector: sorted index vector -> offsets of length dim+1
inline std::vector<NrBoxes> compressIndexVector(const std::vector<NrBoxes>& idx, NrBoxes dim) {
  std::vector<NrBoxes> temp(size_t(dim) + 1, 0);
  size_t i = 0;
  while (i < idx.size()) {
    size_t j = i;
    while (j < idx.size() && idx[j] == idx[i]) ++j;
    if (size_t(idx[i]) + 1 < temp.size()) temp[idx[i] + 1] = NrBoxes(j - i);
    i = j;
  }
  std::vector<NrBoxes> out(temp.size());
  NrBoxes s = 0;
  for (size_t k = 0; k < temp.size(); ++k) { s += temp[k]; out[k] = s; }
  return out;
}

// ---------------------------------------------------------------------------------------------
// The tree: leaf-only sorted SoA (ImplicitBoxTree.h:607-639)
// ---------------------------------------------------------------------------------------------
template <int DIM>
struct PointSet {
  const Grid<DIM>* grid;
  const Map<DIM>* map;
  std::vector<Box> boxes;  // source boxes (already reduced to the iterator's depth)
  Flags flagsToSet = NONE;
  uint64_t size() const { return uint64_t(boxes.size()) * grid->pointsPerBox(); }
};

template <int DIM>
class Tree {
 public:
  Geometry<DIM> geo;
  std::vector<uint64_t> bits;
  std::vector<Depth> depthV;
  std::vector<Flags> flagsV;
  bool evalTwice = false;        // mirror the reference's second image evaluation per search
  uint64_t chunkPoints = ~0ull;  // search buffer size (ImplicitBoxTree.hpp:2643-2680 chunking)

  Tree() { insertSurroundingBoxIfEmpty(); }

  NrBoxes size() const { return NrBoxes(bits.size()); }
  Box box(size_t i) const { return {bits[i], depthV[i], flagsV[i]}; }
  void setBox(size_t i, const Box& b) { bits[i] = b.bits; depthV[i] = b.depth; flagsV[i] = b.flags; }
  void resize(size_t n) { bits.resize(n); depthV.resize(n); flagsV.resize(n); }

  // ImplicitBoxTree.hpp:1959-1970
  void insertSurroundingBoxIfEmpty() {
    if (bits.empty()) { resize(1); setBox(0, surroundingBox()); }
  }

  // ImplicitBoxTree.hpp:1264-1271
  Depth depth() const {
    Depth m = 0;
    for (Depth d : depthV) m = std::max(m, d);
    return m;
  }

  // ImplicitBoxTree.hpp:589-609
  bool areAllLeavesOnDepth(Depth d) const {
    if (d == kLeafDepth) return true;
    if (d == kUnsubDepth) return false;
    for (Depth x : depthV) if (x != d) return false;
    return true;
  }

  // Functors.hpp:337-377 IsUnsubdividableFunctor over (box, pred, succ) of
  // ImplicitBoxTree.hpp:2023-2058 isUnsubdividable_begin
  bool isUnsubdividable(size_t i, Flags f) const {
    size_t n = bits.size();
    Box b = box(i), pr = box(i > 0 ? i - 1 : 0), su = box(std::min(i + 1, n - 1));
    bool res = areFlagsSet(b.flags, f) && b.depth > 0;
    if (res) {
      const Box& nb = isRightHalf(b) ? pr : su;
      res = isSameBox(decrementDepthBy(b, 1), decrementDepthBy(nb, 1))
                ? areFlagsSet(nb.flags, f)
                : !hasSamePath(decrementDepthBy(b, 1), nb);
    }
    return res;
  }

  // Functors.hpp:278-285 IsUnequalToPredecessorInDepthFunctor over
  // ImplicitBoxTree.hpp:1990-2021 isNewBoxInDepth_begin
  bool isNewBoxInDepth(size_t i, Depth d) const {
    Box b = decrementDepthTo(box(i), d);
    Box p = decrementDepthTo(box(i > 0 ? i - 1 : 0), d);
    return hasAtLeastDepth(b, d) && (i == 0 || !isSameBox(b, p));
  }

  // ImplicitBoxTree.hpp:1060-1095 count
  NrBoxes count(Depth d) const {
    if (d == 0) return 1;
    if (d == kLeafDepth || areAllLeavesOnDepth(d)) return size();
    NrBoxes c = 0;
    if (d == kUnsubDepth) {
      for (size_t i = 0; i < bits.size(); ++i) c += isNewBoxInDepth(i, d) && isUnsubdividable(i, NONE);
      return c;
    }
    if (d <= depth()) {
      for (size_t i = 0; i < bits.size(); ++i) c += isNewBoxInDepth(i, d);
      return c;
    }
    return 0;
  }

  // ImplicitBoxTree.hpp:1273-1283 enumerate: {count(0), ..., count(depth())}
  std::vector<NrBoxes> enumerate() const {
    std::vector<NrBoxes> r;
    for (int d = 0; d <= int(depth()); ++d) r.push_back(count(Depth(d)));
    return r;
  }

  // ImplicitBoxTree.hpp:611-775 boxes(depth): {centre, radius, flags} of every box at depth d.
  // The root carries the AND of all leaf flags; at the leaves (or a depth all leaves share) each
  // leaf with its flags; at a depth between root and leaves or the unsubdividable depth the runs
  // of leaves that truncate to the same box, in leaf order, flags ANDed over the run
  // (reduce_by_key / inclusive_scan_by_key + unique), leaves shallower than d (or not
  // unsubdividable) left out; beyond the deepest leaf nothing.
  std::vector<double> boxes(Depth d) const {
    std::vector<double> out;
    auto put = [&](const Box& b, Flags f) {
      double c[DIM], r[DIM];
      geo.coords(b, c, r);
      for (int j = 0; j < DIM; ++j) out.push_back(c[j]);
      for (int j = 0; j < DIM; ++j) out.push_back(r[j]);
      out.push_back(double(f));
    };
    if (d == 0) {
      Flags f = ALL;
      for (Flags x : flagsV) f = Flags(f & x);
      for (int j = 0; j < DIM; ++j) out.push_back(geo.center[j]);
      for (int j = 0; j < DIM; ++j) out.push_back(geo.radius[j]);
      out.push_back(double(f));
      return out;
    }
    if (d == kLeafDepth || areAllLeavesOnDepth(d)) {
      for (size_t i = 0; i < bits.size(); ++i) put(box(i), flagsV[i]);
      return out;
    }
    if (!depthExists(d)) return out;
    const size_t n = bits.size();
    for (size_t i = 0; i < n;) {
      if (!inDepth(i, d)) {
        ++i;
        continue;
      }
      const Box b = decrementDepthTo(box(i), d);
      Flags f = flagsV[i];
      size_t k = i + 1;
      while (k < n && inDepth(k, d) && isSameBox(decrementDepthTo(box(k), d), b)) f = Flags(f & flagsV[k++]);
      put(b, f);
      i = k;
    }
    return out;
  }

  // Depths other than the root, the leaves and a depth every leaf has: a depth d between them
  // (0 < d <= depth()) or the unsubdividable depth. The boxes at such a depth are the leaves'
  // ancestors at d (parents that could be merged, at the unsubdividable depth).
  bool isGeneralDepth(Depth d) const { return d != 0 && d != kLeafDepth && !areAllLeavesOnDepth(d); }
  bool depthExists(Depth d) const { return d == kUnsubDepth || d <= depth(); }
  // leaf i belongs to a box at depth d: hasAtLeastDepth, or unsubdividable at the unsubdividable depth
  bool inDepth(size_t i, Depth d) const {
    return d == kUnsubDepth ? isUnsubdividable(i, NONE) : hasAtLeastDepth(box(i), d);
  }
  // leaf i is the first leaf of a box at depth d (ImplicitBoxTree.hpp:1990-2021, and with
  // isUnsubdividable at the unsubdividable depth)
  bool startsBoxInDepth(size_t i, Depth d) const {
    return isNewBoxInDepth(i, d) && (d != kUnsubDepth || isUnsubdividable(i, NONE));
  }

  // ImplicitBoxTree.hpp:1383-1473 initializeIndicesInDepth: index of the depth-d box of every
  // leaf (inclusive count of box starts, minus one where positive)
  std::vector<NrBoxes> indicesInDepth(Depth d) const {
    std::vector<NrBoxes> idx(bits.size(), 0);
    if (d == 0) return idx;
    if (d == kLeafDepth || areAllLeavesOnDepth(d)) {
      std::iota(idx.begin(), idx.end(), NrBoxes(0));
      return idx;
    }
    NrBoxes acc = 0;
    for (size_t i = 0; i < idx.size(); ++i) {
      acc += startsBoxInDepth(i, d) ? 1 : 0;
      idx[i] = acc > 0 ? acc - 1 : 0;
    }
    return idx;
  }

  // ImplicitBoxTree.hpp:1097-1214 countIfFlags at a general depth: the combination of the flags
  // of every leaf of a box (reverse scan over runs of equal (index, inDepth)) tested at the box's
  // first leaf
  template <class Pred, class Comb>
  NrBoxes countIfFlagsInDepth(const uint8_t* stencil, Depth d, Pred pred, Comb comb,
                              const std::vector<NrBoxes>& idx) const {
    if (!depthExists(d)) return 0;
    const size_t n = bits.size();
    std::vector<Flags> acc(n);
    std::vector<uint8_t> in(n);
    for (size_t i = 0; i < n; ++i) in[i] = inDepth(i, d);
    for (size_t i = n; i-- > 0;) {
      const bool same = i + 1 < n && idx[i + 1] == idx[i] && in[i + 1] == in[i];
      acc[i] = same ? comb(acc[i + 1], flagsV[i]) : flagsV[i];
    }
    NrBoxes c = 0;
    for (size_t i = 0; i < n; ++i) c += (startsBoxInDepth(i, d) && pred(acc[i]) && stencil[idx[i]]) ? 1 : 0;
    return c;
  }

  // countIfFlags (ImplicitBoxTree.hpp:1097-1214) fused into the stencil flag operations
  // setFlagsByStencil :2803-2881, unsetFlagsByStencil :3334-3411, changeFlagsByStencil :912-1058
  template <class Pred, class Comb, class Apply>
  NrBoxes flagsByStencil(const uint8_t* stencil, Depth d, Pred pred, Comb comb, Apply apply) {
    NrBoxes n = 0;
    if (isGeneralDepth(d)) {  // set on every leaf of a selected box (:2848-2878)
      if (!depthExists(d)) return 0;
      const std::vector<NrBoxes> idx = indicesInDepth(d);
      n = countIfFlagsInDepth(stencil, d, pred, comb, idx);
      std::vector<uint8_t> sel(bits.size());
      for (size_t i = 0; i < sel.size(); ++i) sel[i] = inDepth(i, d) && stencil[idx[i]];
      for (size_t i = 0; i < sel.size(); ++i) if (sel[i]) flagsV[i] = apply(flagsV[i]);
      return n;
    }
    if (d == 0) {
      Flags acc = flagsV.front();
      for (Flags f : flagsV) acc = comb(acc, f);
      n = (stencil[0] && pred(acc)) ? 1 : 0;
      if (stencil[0]) for (auto& f : flagsV) f = apply(f);
    } else {
      for (size_t i = 0; i < flagsV.size(); ++i) n += (pred(flagsV[i]) && stencil[i]) ? 1 : 0;
      for (size_t i = 0; i < flagsV.size(); ++i) if (stencil[i]) flagsV[i] = apply(flagsV[i]);
    }
    return n;
  }

  NrBoxes setFlagsByStencil(const uint8_t* st, Flags f, Depth d) {
    return flagsByStencil(st, d, [f](Flags x) { return !areFlagsSet(x, f); },
                          [](Flags a, Flags b) { return Flags(a & b); }, [f](Flags x) { return setF(x, f); });
  }
  NrBoxes unsetFlagsByStencil(const uint8_t* st, Flags f, Depth d) {
    return flagsByStencil(st, d, [f](Flags x) { return isAnyFlagSet(x, f); },
                          [](Flags a, Flags b) { return Flags(a | b); }, [f](Flags x) { return unsetF(x, f); });
  }
  NrBoxes changeFlagsByStencil(const uint8_t* st, Flags fu, Flags fs, Depth d) {
    return flagsByStencil(st, d, [fu](Flags x) { return areFlagsSet(x, fu); },
                          [](Flags a, Flags b) { return Flags(a & b); },
                          [fu, fs](Flags x) { return changeF(x, fu, fs); });
  }

  std::vector<uint8_t> allTrue() const { return std::vector<uint8_t>(std::max<size_t>(bits.size(), 1), 1); }
  NrBoxes setFlags(Flags f, Depth d) { auto s = allTrue(); return setFlagsByStencil(s.data(), f, d); }
  NrBoxes unsetFlags(Flags f, Depth d) { auto s = allTrue(); return unsetFlagsByStencil(s.data(), f, d); }
  NrBoxes changeFlags(Flags fu, Flags fs, Depth d) {
    auto s = allTrue();
    return changeFlagsByStencil(s.data(), fu, fs, d);
  }

  // setFlagsBySearch :2716-2801, unsetFlagsBySearch :3247-3332, changeFlagsBySearch :822-910
  // searchVector is sorted and unique; a trailing invalid entry is ignored.
  template <class Pred, class Apply, class AllPred>
  NrBoxes flagsBySearch(const std::vector<NrBoxes>& sv, Depth d, Pred pred, Apply apply, AllPred rootPred) {
    size_t m = sv.size() - ((!sv.empty() && sv.back() == kInvalid) ? 1 : 0);
    if (d == 0) {
      bool hit = false;
      for (size_t i = 0; i < m; ++i) hit = hit || sv[i] == 0;
      if (hit && rootPred()) {
        for (auto& f : flagsV) f = apply(f);
        return 1;
      }
      return 0;
    }
    NrBoxes n = 0;
    for (size_t i = 0; i < m; ++i) n += pred(flagsV[sv[i]]) ? 1 : 0;
    for (size_t i = 0; i < m; ++i) flagsV[sv[i]] = apply(flagsV[sv[i]]);
    return n;
  }

  NrBoxes setFlagsBySearch(const std::vector<NrBoxes>& sv, Flags f, Depth d) {
    if (isGeneralDepth(d)) {  // :2780-2795 a stencil over the boxes at depth d
      if (!depthExists(d)) return 0;
      auto st = searchStencil(sv, d);
      return setFlagsByStencil(st.data(), f, d);
    }
    return flagsBySearch(sv, d, [f](Flags x) { return !areFlagsSet(x, f); }, [f](Flags x) { return setF(x, f); },
                         [&]() { for (Flags x : flagsV) if (!areFlagsSet(x, f)) return true; return false; });
  }
  NrBoxes unsetFlagsBySearch(const std::vector<NrBoxes>& sv, Flags f, Depth d) {
    if (isGeneralDepth(d)) {  // :2780-2795 a stencil over the boxes at depth d
      if (!depthExists(d)) return 0;
      auto st = searchStencil(sv, d);
      return unsetFlagsByStencil(st.data(), f, d);
    }
    return flagsBySearch(sv, d, [f](Flags x) { return isAnyFlagSet(x, f); }, [f](Flags x) { return unsetF(x, f); },
                         [&]() { for (Flags x : flagsV) if (isAnyFlagSet(x, f)) return true; return false; });
  }
  NrBoxes changeFlagsBySearch(const std::vector<NrBoxes>& sv, Flags fu, Flags fs, Depth d) {
    if (isGeneralDepth(d)) {
      if (!depthExists(d)) return 0;
      auto st = searchStencil(sv, d);
      return changeFlagsByStencil(st.data(), fu, fs, d);
    }
    return flagsBySearch(sv, d, [fu](Flags x) { return areFlagsSet(x, fu); },
                         [fu, fs](Flags x) { return changeF(x, fu, fs); },
                         [&]() { for (Flags x : flagsV) if (!areFlagsSet(x, fu)) return false; return true; });
  }

  std::vector<uint8_t> searchStencil(const std::vector<NrBoxes>& sv, Depth d) const {
    std::vector<uint8_t> st(std::max<NrBoxes>(count(d), 1), 0);
    for (NrBoxes r : sv) if (r != kInvalid) st[r] = 1;
    return st;
  }

  // ImplicitBoxTree.hpp:2343-2371 remove: keep leaves having all of flags (stable)
  NrBoxes remove(Flags f) {
    size_t n = bits.size(), w = 0;
    for (size_t i = 0; i < n; ++i)
      if (areFlagsSet(flagsV[i], f)) { setBox(w, box(i)); ++w; }
    resize(w);
    NrBoxes res = NrBoxes(n - w);
    insertSurroundingBoxIfEmpty();
    return res;
  }

  // ImplicitBoxTree.hpp:1215-1262 deleteDepth: every leaf truncated to depth d-1, equal boxes
  // merged into their first with the AND of their flags (d = 0: the surrounding box alone);
  // returns count(d) before (1 for d = 0, 0 for the leaf, unsubdividable or too deep a depth)
  NrBoxes deleteDepth(Depth d) {
    if (d == 0) {
      resize(0);
      insertSurroundingBoxIfEmpty();
      return 1;
    }
    if (d == kLeafDepth || d == kUnsubDepth || d > depth()) return 0;
    const NrBoxes nb = count(d);
    std::vector<Box> out;
    for (size_t i = 0; i < bits.size(); ++i) {
      const Box b = decrementDepthTo(box(i), Depth(d - 1));
      if (!out.empty() && isSameBox(out.back(), b)) out.back().flags = Flags(out.back().flags & b.flags);
      else out.push_back(b);
    }
    resize(out.size());
    for (size_t i = 0; i < out.size(); ++i) setBox(i, out[i]);
    return nb;
  }

  // ImplicitBoxTree.hpp:1613-1878 insertByIterators on the point boxes pts (maximal depth, or
  // invalid): returns isToInsert per point.
  //  - depth 0: nothing inserted; flagsForHit set on every leaf if some point is valid.
  //  - leaf depth: a point is inserted if no leaf holds it (and it is valid) or the leaf that
  //    holds it is shallower than the maximal depth; every other hit leaf gets flagsForHit. An
  //    inserted box takes the depth one below its deepest mutual ancestor with the leaves on
  //    either side of its lower bound (AdaptDepthOfFirstBoxFunctor, Functors.hpp:379-389).
  //  - a depth d <= depth(): a valid point is inserted if no box at d holds it; hit boxes at d get
  //    flagsForHit; inserted boxes are truncated to d.
  //  - a depth below depth(): every valid point is inserted at d.
  // The new boxes (flagsForInserted set, duplicates dropped) join the leaves; after sorting, each
  // maximal run of consecutive boxes on one path (hasSamePath) becomes one box with the largest
  // pattern and depth of the run and the OR of its flags (inclusive_scan_by_key forwards and
  // backwards), and consecutive boxes on one path then collapse into the first (unique,
  // comparing neighbours).
  std::vector<uint8_t> insertBoxes(const std::vector<Box>& pts, Depth d, Flags fIns, Flags fHit) {
    const size_t P = pts.size();
    std::vector<uint8_t> res(P, 0);
    if (d == 0) {
      bool anyValid = false;
      for (const Box& p : pts) anyValid = anyValid || !isInvalid(p);
      if (fHit != NONE && anyValid)
        for (auto& f : flagsV) f = setF(f, fHit);
      return res;
    }
    if (!(d == kLeafDepth || d <= kMaxDepth)) return res;
    const bool doSearch = d == kLeafDepth || d <= depth();
    std::vector<NrBoxes> sv(P, kInvalid);
    if (doSearch) {
      std::vector<NrBoxes> idx;
      if (isGeneralDepth(d)) idx = indicesInDepth(d);
      for (size_t i = 0; i < P; ++i) sv[i] = searchOne(pts[i], d, &idx);
    }
    if (d == kLeafDepth) {
      std::vector<NrBoxes> hits;
      for (size_t i = 0; i < P; ++i) {
        res[i] = sv[i] == kInvalid ? !isInvalid(pts[i]) : (depthV[sv[i]] < kMaxDepth);
        hits.push_back(res[i] ? kInvalid : sv[i]);
      }
      std::sort(hits.begin(), hits.end());
      hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
      setFlagsBySearch(hits, fHit, kLeafDepth);
    } else if (d <= depth()) {
      for (size_t i = 0; i < P; ++i) res[i] = sv[i] == kInvalid && !isInvalid(pts[i]);
      std::vector<NrBoxes> hits(sv);
      std::sort(hits.begin(), hits.end());
      hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
      setFlagsBySearch(hits, fHit, d);
    } else {
      for (size_t i = 0; i < P; ++i) res[i] = !isInvalid(pts[i]);
    }
    const size_t n = bits.size();
    std::vector<Box> add;
    for (size_t i = 0; i < P; ++i) {
      if (!res[i]) continue;
      Box b = pts[i];
      if (d == kLeafDepth) {
        const size_t lb = lowerBound(b);
        const Box prev = box(lb > 0 ? lb - 1 : 0), next = box(std::min(lb, n - 1));
        b = changeDepthTo(b, Depth(std::max(mutualAncestorDepth(b, prev), mutualAncestorDepth(b, next)) + 1));
      } else {
        b = decrementDepthTo(b, d);
      }
      add.push_back(b);
    }
    std::stable_sort(add.begin(), add.end(), isStrictlyPreceding);
    add.erase(std::unique(add.begin(), add.end(), isSameBox), add.end());
    for (Box& b : add) b.flags = setF(b.flags, fIns);
    for (const Box& b : add) {
      bits.push_back(b.bits);
      depthV.push_back(b.depth);
      flagsV.push_back(b.flags);
    }
    sortBoxes();
    mergeSamePath();
    return res;
  }

  void mergeSamePath() {
    const size_t n = bits.size();
    if (n == 0) return;
    std::vector<Box> rep;
    std::vector<size_t> seg(n);
    for (size_t i = 0; i < n; ++i) {
      const Box b = box(i);
      if (i == 0 || !hasSamePath(box(i - 1), b)) {
        rep.push_back(b);
      } else {
        Box& r = rep.back();
        r.bits = std::max(r.bits, b.bits);
        r.depth = std::max(r.depth, b.depth);
        r.flags = Flags(r.flags | b.flags);
      }
      seg[i] = rep.size() - 1;
    }
    std::vector<Box> out;
    for (size_t t = 0; t < rep.size(); ++t)
      if (t == 0 || !hasSamePath(rep[t - 1], rep[t])) out.push_back(rep[t]);
    resize(out.size());
    for (size_t i = 0; i < out.size(); ++i) setBox(i, out[i]);
  }

  // ImplicitBoxTree.hpp:1880-1892 / 1894-1918 insertByIterators over a point set, insertByPoints
  std::vector<uint8_t> insertByIterators(const PointSet<DIM>& ps, Depth d, Flags fIns, Flags fHit) {
    std::vector<Box> pts(ps.size());
    for (uint64_t i = 0; i < ps.size(); ++i) pts[i] = pointBox(ps, i);
    return insertBoxes(pts, d, fIns, fHit);
  }
  std::vector<uint8_t> insertByPoints(const double* p, uint64_t n, Depth d, Flags fIns, Flags fHit) {
    std::vector<Box> pts(n);
    for (uint64_t i = 0; i < n; ++i) pts[i] = geo.boxFromPoint(p + i * DIM, fIns);
    return insertBoxes(pts, d, fIns, fHit);
  }

  // helpFunctions.hpp:40-98 sortBoxes (ordering = isStrictlyPreceding)
  void sortBoxes() {
    std::vector<Box> v(bits.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = box(i);
    std::stable_sort(v.begin(), v.end(), isStrictlyPreceding);
    for (size_t i = 0; i < v.size(); ++i) setBox(i, v[i]);
  }

  // ImplicitBoxTree.hpp:2883-2930 subdivide; returns #leaves with flags set
  NrBoxes subdivide(Flags f) {
    NrBoxes res = 0;
    for (Flags x : flagsV) res += areFlagsSet(x, f) ? 1 : 0;
    size_t n = bits.size();
    std::vector<Box> add;
    for (size_t i = 0; i < n; ++i) {
      Box b = box(i);
      if (areFlagsSet(b.flags, f) && b.depth < kMaxDepth) {  // Functors.hpp:326-335
        Box l = incrementDepthBy(b, 1);
        setBox(i, l);
        add.push_back(computeOtherHalf(l));
      }
    }
    for (const Box& b : add) {
      resize(bits.size() + 1);
      setBox(bits.size() - 1, b);
    }
    sortBoxes();
    return res;
  }

  // ImplicitBoxTree.hpp:3413-3461 unsubdivide
  NrBoxes unsubdivide(Flags f) {
    size_t n = bits.size();
    std::vector<uint8_t> un(n);
    NrBoxes res = 0;
    for (size_t i = 0; i < n; ++i) { un[i] = isUnsubdividable(i, f); res += un[i]; }
    for (size_t i = 0; i < n; ++i) if (un[i]) setBox(i, decrementDepthBy(box(i), 1));
    // reverse segmented AND of flags over equal boxes, keep first of each run
    std::vector<Box> out;
    size_t i = 0;
    while (i < n) {
      size_t j = i;
      Flags acc = flagsV[i];
      while (j + 1 < n && isSameBox(box(j + 1), box(i))) { ++j; acc &= flagsV[j]; }
      Box b = box(i);
      b.flags = acc;
      out.push_back(b);
      i = j + 1;
    }
    res -= NrBoxes(n - out.size());
    resize(out.size());
    for (size_t k = 0; k < out.size(); ++k) setBox(k, out[k]);
    return res;
  }

  // Point iterators: ImplicitBoxTree.hpp:2082-2120 (root), 2148-2196 (leaves), 2275-2341 (by
  // flags, with initializeLeafIndicesFromDepth :1475-1611 at leaf depth / depth 0)
  PointSet<DIM> makePoints(const Grid<DIM>& grid, const Map<DIM>& map, Depth d, Flags check) const {
    PointSet<DIM> ps{&grid, &map, {}, NONE};
    if (d == 0) {
      bool ok = true;
      if (check != NONE) for (Flags x : flagsV) ok = ok && areFlagsSet(x, check);
      if (ok) ps.boxes.push_back(surroundingBox());
    } else if (d == kLeafDepth || areAllLeavesOnDepth(d)) {
      for (size_t i = 0; i < bits.size(); ++i)
        if (check == NONE || areFlagsSet(flagsV[i], check)) ps.boxes.push_back(box(i));
    } else if (depthExists(d)) {  // :1528-1601 the first leaf of every box at depth d
      const size_t n = bits.size();
      std::vector<uint8_t> ok(n, 1);
      if (check != NONE)  // AND of the check over runs of leaves with the same box at depth d
        for (size_t i = n; i-- > 0;) {
          const bool same = i + 1 < n && isSameBox(decrementDepthTo(box(i + 1), d), decrementDepthTo(box(i), d));
          ok[i] = areFlagsSet(flagsV[i], check) && (same ? ok[i + 1] : 1);
        }
      for (size_t i = 0; i < n; ++i)
        if (startsBoxInDepth(i, d) && ok[i]) ps.boxes.push_back(decrementDepthTo(box(i), d));
    }
    return ps;
  }

  // Functors.hpp:476-511 GridPointsFunctor::operator()
  Box pointBox(const PointSet<DIM>& ps, uint64_t i) const {
    double c[3 * DIM];
    const Box& b = ps.boxes[i / ps.grid->pointsPerBox()];
    geo.coords(b, c, c + DIM);
    (*ps.grid)(i, c + 2 * DIM);
    for (int j = 0; j < DIM; ++j) c[j] += c[j + 2 * DIM] * c[j + DIM];
    (*ps.map)(c, c + DIM);
    return geo.boxFromPoint(c + DIM, ps.flagsToSet);
  }

  // lower_bound with isStrictlyPreceding over the leaves (ImplicitBoxTree.hpp:2397-2401)
  NrBoxes lowerBound(const Box& p) const {
    size_t lo = 0, hi = bits.size();
    while (lo < hi) {
      size_t mid = (lo + hi) / 2;
      if (isStrictlyPreceding(box(mid), p)) lo = mid + 1; else hi = mid;
    }
    return NrBoxes(lo);
  }

  // Functors.hpp:287-315 ComputeSearchIndexFunctor at leaf depth
  NrBoxes searchIndex(NrBoxes s, const Box& p) const {
    NrBoxes n = size();
    Box b1 = box(s > 0 ? s - 1 : 0);
    Box b2 = box(std::min<NrBoxes>(s, n - 1));
    if (p.depth >= b1.depth && hasSamePath(p, b1)) return s > 0 ? s - 1 : 0;
    if (p.depth >= b2.depth && hasSamePath(p, b2)) return s;
    return kInvalid;
  }

  // ComputeSearchIndexFunctor with the two neighbours decremented to depth d and their inDepth
  // flags, then the leaf's box index at d (ImplicitBoxTree.hpp:2433-2566)
  NrBoxes searchIndexInDepth(NrBoxes s, const Box& p, Depth d, const std::vector<NrBoxes>& idx) const {
    NrBoxes n = size();
    const NrBoxes i1 = s > 0 ? s - 1 : 0, i2 = std::min<NrBoxes>(s, n - 1);
    const Box b1 = decrementDepthTo(box(i1), d), b2 = decrementDepthTo(box(i2), d);
    NrBoxes r = kInvalid;
    if (inDepth(i1, d) && p.depth >= b1.depth && hasSamePath(p, b1)) r = i1;
    else if (inDepth(i2, d) && p.depth >= b2.depth && hasSamePath(p, b2)) r = s;
    return r == kInvalid ? kInvalid : idx[r];
  }

  NrBoxes searchOne(const Box& p, Depth d, const std::vector<NrBoxes>* idx = nullptr) const {
    if (d == 0) return isInvalid(p) ? kInvalid : 0;
    if (isGeneralDepth(d)) return depthExists(d) ? searchIndexInDepth(lowerBound(p), p, d, *idx) : kInvalid;
    return searchIndex(lowerBound(p), p);
  }

  // ImplicitBoxTree.hpp:2373-2432 searchByIterators (depth 0 and leaf depth)
  std::vector<NrBoxes> search(const PointSet<DIM>& ps, Depth d) const {
    uint64_t P = ps.size();
    std::vector<NrBoxes> out(P);
    const bool general = isGeneralDepth(d);
    std::vector<NrBoxes> idx;
    if (general) idx = indicesInDepth(d);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(P); ++i) {
      Box p = pointBox(ps, uint64_t(i));
      if (evalTwice) {
        NrBoxes s = (d == 0) ? 0 : lowerBound(p);
        Box p2 = pointBox(ps, uint64_t(i));
        if (d == 0) out[i] = isInvalid(p2) ? kInvalid : 0;
        else if (general) out[i] = depthExists(d) ? searchIndexInDepth(s, p2, d, idx) : kInvalid;
        else out[i] = searchIndex(s, p2);
      } else {
        out[i] = searchOne(p, d, &idx);
      }
    }
    return out;
  }

  std::vector<NrBoxes> searchPoints(const double* pts, uint64_t n, Depth d) const {
    std::vector<NrBoxes> out(n);
    std::vector<NrBoxes> idx;
    if (isGeneralDepth(d)) idx = indicesInDepth(d);
    for (uint64_t i = 0; i < n; ++i) out[i] = searchOne(geo.boxFromPoint(pts + i * DIM, NONE), d, &idx);
    return out;
  }

  // ImplicitBoxTree.hpp:2622-2681 searchCompacted: sorted unique hit indices, chunked
  std::vector<NrBoxes> searchCompacted(const PointSet<DIM>& ps, Depth d) const {
    std::vector<NrBoxes> res;
    if (d == 0) {
      bool allInvalid = true;
      for (uint64_t i = 0; i < ps.size(); ++i) allInvalid = allInvalid && isInvalid(pointBox(ps, i));
      if (!allInvalid) res.push_back(0);
      return res;
    }
    std::vector<NrBoxes> all = search(ps, d);
    NrBoxes nBoxes = count(d);
    uint64_t P = all.size(), start = 0;
    while (start < P && res.size() < nBoxes) {
      uint64_t m = std::min<uint64_t>(chunkPoints, P - start);
      res.insert(res.end(), all.begin() + start, all.begin() + start + m);
      std::sort(res.begin(), res.end());
      res.erase(std::unique(res.begin(), res.end()), res.end());
      if (!res.empty() && res.back() == kInvalid) res.pop_back();
      start += m;
    }
    return res;
  }

  // ImplicitBoxTree.hpp:2932-3159 transitionMatrix (COO; invalid targets dropped or, with
  // useInvalid, relabelled to column nColumns after its increment, as the reference does)
  Coo transitionMatrix(const PointSet<DIM>& ps, Depth d, bool useAbs, bool rowMajor, bool useInvalid) const {
    std::vector<NrBoxes> col = search(ps, d);
    uint64_t k = ps.grid->pointsPerBox();
    std::vector<uint64_t> key(col.size());
    for (uint64_t i = 0; i < col.size(); ++i) key[i] = (uint64_t(NrBoxes(i / k)) << 32) | col[i];
    std::sort(key.begin(), key.end());
    std::vector<uint64_t> uk;
    std::vector<uint64_t> cnt;
    for (uint64_t i = 0; i < key.size();) {
      uint64_t j = i;
      while (j < key.size() && key[j] == key[i]) ++j;
      uk.push_back(key[i]);
      cnt.push_back(j - i);
      i = j;
    }
    Coo m;
    m.nRows = m.nCols = count(d);
    bool anyInvalid = false;
    for (uint64_t x : uk) anyInvalid = anyInvalid || NrBoxes(x) == kInvalid;
    if (anyInvalid && useInvalid) m.nCols += 1;
    std::vector<std::pair<uint64_t, uint64_t>> e;  // (sort key, count)
    for (size_t i = 0; i < uk.size(); ++i) {
      NrBoxes r = NrBoxes(uk[i] >> 32), c = NrBoxes(uk[i]);
      if (c == kInvalid) {
        if (!useInvalid) continue;
        if (rowMajor) c = m.nCols;  // replace_if before any re-sort
      }
      uint64_t sk = rowMajor ? ((uint64_t(r) << 32) | c) : ((uint64_t(c) << 32) | r);
      e.push_back({sk, cnt[i]});
    }
    if (!rowMajor) std::sort(e.begin(), e.end());
    for (auto& x : e) {
      NrBoxes a = NrBoxes(x.first >> 32), b = NrBoxes(x.first);
      NrBoxes r = rowMajor ? a : b, c = rowMajor ? b : a;
      if (!rowMajor && c == kInvalid) c = m.nCols;  // fill after the column-major sort
      m.rows.push_back(r);
      m.cols.push_back(c);
      m.vals.push_back(useAbs ? double(x.second) : double(x.second) / double(k));
    }
    return m;
  }
};

// ---------------------------------------------------------------------------------------------
// Invariant measure (build-defined; the reference leaves it to MATLAB, subdivisionByNim.m:30-59).
// Power iteration on unnormalised iterates, row-vector convention, P given as COO:
//   z_0 = mu_0 (1/n or the caller's vector), nu_0 = 1
//   z_t[j] = c_t * sum_{i in column j, ascending} z_{t-1}[i] * P_ij        (sum from 0.0)
//   nu_t = S(|z_t|),  r_t = S(|z_t[j] / nu_t - z_{t-1}[j] / nu_{t-1}|)
//   stop at the first t with (tol >= 0 and r_t <= tol) or t = maxIter; mu = z_t / nu_t.
// c_t is a power of two that keeps z_t in range: 1, except at t = 16, 24, 32, ... where it is
// 2^-ilogb(nu_{t-8}) when nu_{t-8} is finite, positive and its exponent lies outside [-32, 32].
// A power of two scales z_t and nu_t exactly, so mu_t = z_t / nu_t is that of the normalised
// iteration mu_t = mu_{t-1} P / ||mu_{t-1} P||_1 with the normalisation deferred.
// S is the sum in a fixed order over chunks of 256 columns: within a chunk, each group of 64
// columns by a butterfly (x_l += x_{l^o} for o = 32, 16, ..., 1; entries past n are 0), the four
// group sums added in order to 0.0; over chunks, lane l (0..63) adds chunks l, l+64, ... to 0.0
// in ascending order, and the 64 lane sums combine by the same butterfly.
// ---------------------------------------------------------------------------------------------
inline double butterfly64(double* x) {
  for (int o = 32; o > 0; o >>= 1) {
    double y[64];
    for (int l = 0; l < 64; ++l) y[l] = x[l] + x[l ^ o];
    for (int l = 0; l < 64; ++l) x[l] = y[l];
  }
  return x[0];
}

template <class F>
inline double canonicalSum(size_t n, F&& value) {
  const size_t chunks = (n + 255) / 256;
  double lane[64];
  for (int l = 0; l < 64; ++l) lane[l] = 0.0;
  for (size_t q = 0; q < chunks; ++q) {
    double c = 0.0;
    for (int w = 0; w < 4; ++w) {
      double x[64];
      for (int l = 0; l < 64; ++l) {
        const size_t j = q * 256 + size_t(w) * 64 + size_t(l);
        x[l] = j < n ? value(j) : 0.0;
      }
      c += butterfly64(x);
    }
    lane[q % 64] += c;
  }
  return butterfly64(lane);
}

// the scale c_t of iteration t from the norm nu_{t-8}
inline double piScale(uint64_t t, double nuLag) {
  if (t < 16 || t % 8 != 0) return 1.0;
  if (!(nuLag > 0.0) || !std::isfinite(nuLag)) return 1.0;
  const int e = std::ilogb(nuLag);
  if (e >= -32 && e <= 32) return 1.0;
  return std::ldexp(1.0, -std::max(-1000, std::min(1000, e)));
}

inline int powerIteration(const Coo& P, int maxIter, double tol, std::vector<double>& mu, double* residOut) {
  NrBoxes n = P.nRows;
  std::vector<size_t> order(P.rows.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return P.cols[a] != P.cols[b] ? P.cols[a] < P.cols[b] : P.rows[a] < P.rows[b];
  });
  std::vector<NrBoxes> cptr(size_t(n) + 1, 0);
  for (size_t e : order) if (P.cols[e] < n) ++cptr[P.cols[e] + 1];
  for (NrBoxes j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  std::vector<NrBoxes> ri;
  std::vector<double> va;
  for (size_t e : order) if (P.cols[e] < n) { ri.push_back(P.rows[e]); va.push_back(P.vals[e]); }
  if (mu.size() != n) mu.assign(n, 1.0 / double(n));
  std::vector<double> z = mu, nz(n);
  std::vector<double> nuHist;  // nu_t for t = 0, 1, ...
  nuHist.push_back(1.0);
  int it = 0;
  double resid = -1.0;
  for (it = 0; it < maxIter;) {
    const uint64_t t = uint64_t(it) + 1;
    const double c = piScale(t, t >= 8 ? nuHist[t - 8] : 0.0);
    for (NrBoxes j = 0; j < n; ++j) {
      double s = 0.0;
      for (NrBoxes e = cptr[j]; e < cptr[j + 1]; ++e) s += z[ri[e]] * va[e];
      nz[j] = c != 1.0 ? s * c : s;
    }
    const double nu = canonicalSum(n, [&](size_t j) { return std::fabs(nz[j]); });
    const double nuPrev = nuHist.back();
    const double r = canonicalSum(n, [&](size_t j) { return std::fabs(nz[j] / nu - z[j] / nuPrev); });
    nuHist.push_back(nu);
    z.swap(nz);
    ++it;
    resid = r;
    if (tol >= 0.0 && r <= tol) break;
  }
  if (it > 0)
    for (NrBoxes j = 0; j < n; ++j) mu[j] = z[j] / nuHist.back();
  if (residOut) *residOut = resid;
  return it;
}

// ---------------------------------------------------------------------------------------------
// Energy band of the Hénon-Heiles Hamiltonian (build-defined; not a reference feature)
// H(q, p) = (p1^2 + p2^2)/2 + (q1^2 + q2^2)/2 + q1^2 q2 - q2^3/3 over the box
// [c - r, c + r] in (q1, q2, p1, p2), as an interval in round-to-nearest FP64 in this order:
// squares by sign case, q1^2 q2 as the hull of the four corner products, q2^3 by monotonicity.
// A box meets the band [emin, emax] iff hi >= emin and lo <= emax.
// ---------------------------------------------------------------------------------------------
struct Interval {
  double lo, hi;
};

inline Interval intervalSquare(double a, double b) {
  if (a >= 0.0) return {a * a, b * b};
  if (b <= 0.0) return {b * b, a * a};
  const double x = a * a, y = b * b;
  return {0.0, x > y ? x : y};
}

inline Interval intervalProduct(Interval u, double c, double d) {
  const double p1 = u.lo * c, p2 = u.lo * d, p3 = u.hi * c, p4 = u.hi * d;
  const double lo12 = p1 < p2 ? p1 : p2, lo34 = p3 < p4 ? p3 : p4;
  const double hi12 = p1 > p2 ? p1 : p2, hi34 = p3 > p4 ? p3 : p4;
  return {lo12 < lo34 ? lo12 : lo34, hi12 > hi34 ? hi12 : hi34};
}

inline Interval henonHeilesEnergy(const double* c, const double* r) {
  const double q1a = c[0] - r[0], q1b = c[0] + r[0], q2a = c[1] - r[1], q2b = c[1] + r[1];
  const double p1a = c[2] - r[2], p1b = c[2] + r[2], p2a = c[3] - r[3], p2b = c[3] + r[3];
  const Interval Q1 = intervalSquare(q1a, q1b), Q2 = intervalSquare(q2a, q2b);
  const Interval P1 = intervalSquare(p1a, p1b), P2 = intervalSquare(p2a, p2b);
  const Interval M = intervalProduct(Q1, q2a, q2b);
  const double c3lo = (q2a * q2a) * q2a, c3hi = (q2b * q2b) * q2b;
  const double lo = ((0.5 * (P1.lo + P2.lo) + 0.5 * (Q1.lo + Q2.lo)) + M.lo) - c3hi / 3.0;
  const double hi = ((0.5 * (P1.hi + P2.hi) + 0.5 * (Q1.hi + Q2.hi)) + M.hi) - c3lo / 3.0;
  return {lo, hi};
}

inline std::vector<uint8_t> energyBandStencil(const Tree<4>& T, double emin, double emax) {
  std::vector<uint8_t> st(T.size());
  for (size_t i = 0; i < st.size(); ++i) {
    double c[4], r[4];
    T.geo.coords(T.box(i), c, r);
    const Interval e = henonHeilesEnergy(c, r);
    st[i] = (e.hi >= emin && e.lo <= emax) ? 1 : 0;
  }
  return st;
}

// ---------------------------------------------------------------------------------------------
// Subspace iteration for the m leading eigenvalues (build-defined; the reference leaves
// eigenvalues to MATLAB eigs, subdivisionByNim.m:30-59). Row-vector convention as the power
// iteration: Y = X P with y_j[c] = sum_i x_i[c] P_ij over column j in ascending row order.
//   X_0: column 0 = 1/n, column c > 0 = u(i, c) = (splitmix64(8 i + c) >> 11) 2^-53 * 2 - 1,
//   orthonormalised by Cholesky QR applied twice (Gram G = Y^T Y, G = R^T R, X = Y R^-1
//   column by column: x_c = (y_c - sum_{d<c} R_dc x_d) / R_cc, d ascending).
//   For k = 1, 2, ...: Y = X_{k-1} P, the Ritz matrix H_ab = x_a . y_b and the residual
//   rho_k = ||Y - X_{k-1} H||_F / ||H||_F; stop when tol >= 0 and rho_k <= tol (the result is
//   H and X_{k-1}, k - 1 basis updates) or when k - 1 = maxIter; else X_k = cholQR2(Y).
// Every dot product over the boxes runs in one fixed order: T = 256 min(ceil(n/256), 1024)
// strided partial sums (partial t adds boxes t, t + T, ... to 0.0), each block of 256 partials
// reduced by the wave butterfly and the four wave sums added to 0.0 in order, the block sums
// then strided over 256 lanes (lane l adds blocks l, l + 256, ...) and reduced the same way.
// ||H||_F^2 adds H_ab^2 in row-major order; the two square roots and the quotient follow.
// X is stored column by column (X[c * n + i]).
// ---------------------------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

inline double subspaceStart(uint64_t i, int c) {
  return double(splitmix64(8 * i + uint64_t(c)) >> 11) * 0x1p-53 * 2.0 - 1.0;
}

// 256 partial sums: four waves of 64 by the butterfly, the wave sums added to 0.0 in order
inline double blockSum256(const double* v) {
  double s = 0.0;
  for (int w = 0; w < 4; ++w) {
    double x[64];
    for (int l = 0; l < 64; ++l) x[l] = v[w * 64 + l];
    s += butterfly64(x);
  }
  return s;
}

// sum over i < n of term(i) in the fixed strided order above
template <class F>
inline double stridedSum(size_t n, F&& term) {
  const size_t nb = (n + 255) / 256, nd = std::min<size_t>(std::max<size_t>(nb, 1), 1024), T = nd * 256;
  std::vector<double> part(nd);
  double lanes[256];
  for (size_t b = 0; b < nd; ++b) {
    for (size_t l = 0; l < 256; ++l) {
      double acc = 0.0;
      for (size_t i = b * 256 + l; i < n; i += T) acc += term(i);
      lanes[l] = acc;
    }
    part[b] = blockSum256(lanes);
  }
  for (size_t l = 0; l < 256; ++l) {
    double acc = 0.0;
    for (size_t q = l; q < nd; q += 256) acc += part[q];
    lanes[l] = acc;
  }
  return blockSum256(lanes);
}

// false if the Gram matrix is not positive definite
inline bool cholQR(std::vector<double>& X, NrBoxes n, int m) {
  std::vector<double> G(size_t(m) * m), R(size_t(m) * m, 0.0);
  for (int a = 0; a < m; ++a)
    for (int b = a; b < m; ++b)
      G[a * m + b] = stridedSum(n, [&](size_t i) { return X[size_t(a) * n + i] * X[size_t(b) * n + i]; });
  for (int j = 0; j < m; ++j) {
    double d = G[j * m + j];
    for (int k = 0; k < j; ++k) d -= R[k * m + j] * R[k * m + j];
    if (!(d > 0.0)) return false;
    R[j * m + j] = std::sqrt(d);
    for (int l = j + 1; l < m; ++l) {
      double v = G[j * m + l];
      for (int k = 0; k < j; ++k) v -= R[k * m + j] * R[k * m + l];
      R[j * m + l] = v / R[j * m + j];
    }
  }
  for (NrBoxes i = 0; i < n; ++i)
    for (int c = 0; c < m; ++c) {
      double v = X[size_t(c) * n + i];
      for (int d = 0; d < c; ++d) v -= R[d * m + c] * X[size_t(d) * n + i];
      X[size_t(c) * n + i] = v / R[c * m + c];
    }
  return true;
}

// ||Y - X H||_F / ||H||_F: per box the squares of (y_c - sum_d x_d H_dc) added over c in order
// ||Y_L - X_L H_LL||_F / ||H_LL||_F over the leading `lead` columns (L) of the basis and the
// leading lead x lead block of H: zero when the span of those columns is invariant
inline double subspaceResidual(const std::vector<double>& X, const std::vector<double>& Y,
                               const std::vector<double>& H, NrBoxes n, int m, int lead) {
  const double r2 = stridedSum(n, [&](size_t i) {
    double r = 0.0;
    for (int c = 0; c < lead; ++c) {
      double v = Y[size_t(c) * n + i];
      for (int d = 0; d < lead; ++d) v -= X[size_t(d) * n + i] * H[d * m + c];
      r += v * v;
    }
    return r;
  });
  double h2 = 0.0;
  for (int a = 0; a < lead; ++a)
    for (int b = 0; b < lead; ++b) h2 += H[a * m + b] * H[a * m + b];
  return std::sqrt(r2) / std::sqrt(h2);
}

// the number of basis updates, or -1 if a Gram matrix was not positive definite
inline int subspaceIteration(const Coo& P, int m, int lead, int maxIter, double tol, std::vector<double>& H,
                             std::vector<double>& X, double* residOut) {
  const NrBoxes n = P.nRows;
  std::vector<size_t> order(P.rows.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return P.cols[a] != P.cols[b] ? P.cols[a] < P.cols[b] : P.rows[a] < P.rows[b];
  });
  std::vector<NrBoxes> cptr(size_t(n) + 1, 0);
  for (size_t e : order) if (P.cols[e] < n) ++cptr[P.cols[e] + 1];
  for (NrBoxes j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  std::vector<NrBoxes> ri;
  std::vector<double> va;
  for (size_t e : order) if (P.cols[e] < n) { ri.push_back(P.rows[e]); va.push_back(P.vals[e]); }
  auto spmm = [&](const std::vector<double>& in, std::vector<double>& out) {
    out.assign(size_t(m) * n, 0.0);
    for (NrBoxes j = 0; j < n; ++j)
      for (int c = 0; c < m; ++c) {
        double s = 0.0;
        for (NrBoxes e = cptr[j]; e < cptr[j + 1]; ++e) s += in[size_t(c) * n + ri[e]] * va[e];
        out[size_t(c) * n + j] = s;
      }
  };
  auto ritz = [&](const std::vector<double>& A, const std::vector<double>& B) {
    H.assign(size_t(m) * m, 0.0);
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b)
        H[a * m + b] = stridedSum(n, [&](size_t i) { return A[size_t(a) * n + i] * B[size_t(b) * n + i]; });
  };
  X.assign(size_t(m) * n, 0.0);
  for (NrBoxes i = 0; i < n; ++i)
    for (int c = 0; c < m; ++c) X[size_t(c) * n + i] = c == 0 ? 1.0 / double(n) : subspaceStart(i, c);
  if (!cholQR(X, n, m) || !cholQR(X, n, m)) return -1;
  std::vector<double> Y;
  for (int k = 0;; ++k) {
    spmm(X, Y);
    ritz(X, Y);
    if (tol >= 0.0 || k == maxIter) {
      const double rho = subspaceResidual(X, Y, H, n, m, lead);
      if (residOut) *residOut = rho;
      if (k == maxIter || rho <= tol) return k;
    }
    if (!cholQR(Y, n, m) || !cholQR(Y, n, m)) return -1;
    X.swap(Y);
  }
}

// ---------------------------------------------------------------------------------------------
// Krylov basis of the Krylov-Schur eigensolver (build-defined; the reference leaves eigenvalues
// to MATLAB eigs, subdivisionByNim.m:30-56). The driver (eigs() in b12_amd/api.py) keeps the
// small projected matrix; the basis operations over the boxes are defined here:
//   operator: w = v P, w_j = sum over column j in ascending row order (from 0.0);
//   start: V_0 = v0 / ||v0||, ||v0|| = sqrt(S(v0_i v0_i)), default v0_i = u(i, 1) (subspaceStart);
//   expand(p): w = V_p P; twice: h^(q)_c = S(V_c[i] w_i) for c <= p, then w_i -= V_c[i] h^(q)_c
//     for c = 0..p in order; h_c = h^(1)_c + h^(2)_c; beta = sqrt(S(w_i w_i)); V_{p+1} = w / beta
//     (left as it was when beta = 0);
//   rotate(m, k, q): V_c[i] = sum over d < m of V_d[i] q[d k + c] (from 0.0, d ascending) for c < k,
//     then V_k = V_m.
// S is stridedSum's fixed order. V is stored column by column (V[c * n + i]).
// ---------------------------------------------------------------------------------------------
struct KrylovBasis {
  NrBoxes n = 0;
  int ncv = 0;
  std::vector<NrBoxes> cptr, ri;
  std::vector<double> va, V, w;

  KrylovBasis(const Coo& P, int ncvIn) : n(P.nRows), ncv(ncvIn) {
    std::vector<size_t> order(P.rows.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      return P.cols[a] != P.cols[b] ? P.cols[a] < P.cols[b] : P.rows[a] < P.rows[b];
    });
    cptr.assign(size_t(n) + 1, 0);
    for (size_t e : order) if (P.cols[e] < n) ++cptr[P.cols[e] + 1];
    for (NrBoxes j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
    for (size_t e : order)
      if (P.cols[e] < n) {
        ri.push_back(P.rows[e]);
        va.push_back(P.vals[e]);
      }
    V.assign(size_t(ncv + 1) * n, 0.0);
    w.assign(n, 0.0);
  }
  double* col(int c) { return V.data() + size_t(c) * n; }

  double start(const double* v0) {
    for (NrBoxes i = 0; i < n; ++i) w[i] = v0 ? v0[i] : subspaceStart(i, 1);
    const double beta = std::sqrt(stridedSum(n, [&](size_t i) { return w[i] * w[i]; }));
    if (beta > 0.0)
      for (NrBoxes i = 0; i < n; ++i) col(0)[i] = w[i] / beta;
    return beta;
  }

  double expand(int p, double* h) {
    const double* v = col(p);
    for (NrBoxes j = 0; j < n; ++j) {
      double s = 0.0;
      for (NrBoxes e = cptr[j]; e < cptr[j + 1]; ++e) s += v[ri[e]] * va[e];
      w[j] = s;
    }
    std::vector<double> h1(size_t(p) + 1), h2(size_t(p) + 1);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double>& hq = pass == 0 ? h1 : h2;
      for (int c = 0; c <= p; ++c) {
        const double* x = col(c);
        hq[c] = stridedSum(n, [&](size_t i) { return x[i] * w[i]; });
      }
      for (NrBoxes i = 0; i < n; ++i) {
        double t = w[i];
        for (int c = 0; c <= p; ++c) t -= col(c)[i] * hq[c];
        w[i] = t;
      }
    }
    for (int c = 0; c <= p; ++c) h[c] = h1[c] + h2[c];
    const double beta = std::sqrt(stridedSum(n, [&](size_t i) { return w[i] * w[i]; }));
    if (beta > 0.0)
      for (NrBoxes i = 0; i < n; ++i) col(p + 1)[i] = w[i] / beta;
    return beta;
  }

  void rotate(int m, int k, const double* q) {
    std::vector<double> x(m);
    for (NrBoxes i = 0; i < n; ++i) {
      for (int d = 0; d < m; ++d) x[d] = col(d)[i];
      for (int c = 0; c < k; ++c) {
        double s = 0.0;
        for (int d = 0; d < m; ++d) s += x[d] * q[d * k + c];
        col(c)[i] = s;
      }
    }
    std::copy(col(m), col(m) + n, col(k));
  }
};

// ---------------------------------------------------------------------------------------------
// Strongly connected components of the graph of a matrix (an edge row -> column per entry;
// entries naming a vertex >= nRows, such as the column of points outside the cover, are left
// out): chainRecurrentSet.m:26-39 labels them with MATLAB's graphconncomp. Tarjan's algorithm,
// iterative. Each vertex gets the largest vertex index of its component as label (a canonical
// labelling: the partition alone is what graphconncomp's labels carry).
// ---------------------------------------------------------------------------------------------
inline std::vector<NrBoxes> strongComponents(const Coo& P) {
  const NrBoxes n = P.nRows;
  std::vector<size_t> off(size_t(n) + 1, 0);
  for (size_t e = 0; e < P.rows.size(); ++e)
    if (P.rows[e] < n && P.cols[e] < n) ++off[P.rows[e] + 1];
  for (NrBoxes v = 0; v < n; ++v) off[v + 1] += off[v];
  std::vector<NrBoxes> adj(off[n]);
  std::vector<size_t> fill(off.begin(), off.end() - 1);
  for (size_t e = 0; e < P.rows.size(); ++e)
    if (P.rows[e] < n && P.cols[e] < n) adj[fill[P.rows[e]]++] = P.cols[e];
  const NrBoxes none = 0xFFFFFFFFu;
  std::vector<NrBoxes> index(n, none), low(n, 0), label(n, none), stack;
  std::vector<char> onStack(n, 0);
  std::vector<std::pair<NrBoxes, size_t>> call;  // (vertex, next edge offset)
  NrBoxes counter = 0;
  for (NrBoxes r = 0; r < n; ++r) {
    if (index[r] != none) continue;
    call.push_back({r, off[r]});
    index[r] = low[r] = counter++;
    stack.push_back(r);
    onStack[r] = 1;
    while (!call.empty()) {
      const NrBoxes v = call.back().first;
      size_t& e = call.back().second;
      if (e < off[v + 1]) {
        const NrBoxes w = adj[e++];
        if (index[w] == none) {
          index[w] = low[w] = counter++;
          stack.push_back(w);
          onStack[w] = 1;
          call.push_back({w, off[w]});
        } else if (onStack[w]) {
          low[v] = std::min(low[v], index[w]);
        }
        continue;
      }
      if (low[v] == index[v]) {  // v is the root of a component: pop it
        size_t k = stack.size();
        NrBoxes mx = 0;
        do {
          --k;
          mx = std::max(mx, stack[k]);
        } while (stack[k] != v);
        for (size_t j = k; j < stack.size(); ++j) {
          label[stack[j]] = mx;
          onStack[stack[j]] = 0;
        }
        stack.resize(k);
      }
      call.pop_back();
      if (!call.empty()) {
        const NrBoxes u = call.back().first;
        low[u] = std::min(low[u], low[v]);
      }
    }
  }
  return label;
}

}  // namespace b12cpu
