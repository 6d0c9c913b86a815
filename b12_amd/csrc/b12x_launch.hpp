// Launch helpers shared by the C ABI translation units: array scan, transition-matrix rows,
// stable radix sort.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "b12x_host.hpp"
#include "b12x_kernels.hpp"

namespace b12x {

// Exclusive scan of a device uint32 array in place on a stream; returns the total (synchronises),
// or, with readTotal false, leaves it in *dTotal and returns 0 without synchronising.
inline uint32_t scanArray(hipStream_t st, uint32_t* v, uint64_t n, uint32_t* scratchTotals, uint32_t* dTotal,
                          uint32_t* hTotal, bool readTotal = true) {
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) throw HipFail(B12X_E_ARG, "scan too large");
  const uint32_t nb = ceilDiv(n, kTile);
  ScanArrayOp op{};
  op.v = v;
  hipLaunchKernelGGL(scanCountKernel<ScanArrayOp>, dim3(nb), dim3(kBlock), 0, st, op, uint32_t(n),
                     (const uint32_t*)nullptr, scratchTotals, (uint32_t*)nullptr);
  hipLaunchKernelGGL(scanTotalsKernel, dim3(1), dim3(kBlock), 0, st, scratchTotals, nb, (const uint32_t*)nullptr,
                     dTotal, (const uint32_t*)nullptr, (uint32_t*)nullptr);
  hipLaunchKernelGGL(scanArrayScatterKernel, dim3(nb), dim3(kBlock), 0, st, v, uint32_t(n), scratchTotals);
  HIPCHK(hipGetLastError());
  if (!readTotal) return 0;
  HIPCHK(hipMemcpyAsync(hTotal, dTotal, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return *hTotal;
}

template <int S>
void launchTmRows(const TmRowArgs& a, hipStream_t st) {
  const uint64_t groups = (a.nRows + (kWave / S) - 1) / (kWave / S);
  const uint64_t blocks64 = (groups + (kBlock / kWave) - 1) / (kBlock / kWave);
  const uint32_t blocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(blocks64, 1u << 20)));
  hipLaunchKernelGGL(tmRowsKernel<S>, dim3(blocks), dim3(kBlock), 0, st, a);
}

inline void tmRows(const TmRowArgs& a, hipStream_t st) {
  if (a.k <= 1) launchTmRows<1>(a, st);
  else if (a.k <= 2) launchTmRows<2>(a, st);
  else if (a.k <= 4) launchTmRows<4>(a, st);
  else if (a.k <= 8) launchTmRows<8>(a, st);
  else if (a.k <= 16) launchTmRows<16>(a, st);
  else if (a.k <= 32) launchTmRows<32>(a, st);
  else launchTmRows<64>(a, st);
  HIPCHK(hipGetLastError());
}

// Stable sort of (key, p0, p1) triples by key using ceil(bits/8) LSD passes; with sync false the
// sort is only enqueued on st (its scratch returns to the cache, which fences reuse).
inline void radixSort3(hipStream_t st, uint32_t*& k, uint32_t*& p0, uint32_t*& p1, uint64_t n, int bits,
                uint32_t* hTotal, bool sync = true) {
  if (n <= 1) return;
  uint32_t* k2 = dalloc<uint32_t>(n);
  uint32_t* a2 = dalloc<uint32_t>(n);
  uint32_t* b2 = dalloc<uint32_t>(n);
  const uint32_t nb = ceilDiv(n, kRadixTile);
  const uint64_t nh = uint64_t(nb) * 256;
  uint32_t* hist = dalloc<uint32_t>(nh);
  uint32_t* tt = dalloc<uint32_t>(nh / kTile + 16);
  uint32_t* dTot = dalloc<uint32_t>(1);
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(radixHistKernel, dim3(nb), dim3(kBlock), 0, st, k, n, shift, hist, nb);
    scanArray(st, hist, nh, tt, dTot, hTotal, false);  // the total is n: no read-back
    hipLaunchKernelGGL(radixScatterKernel, dim3(nb), dim3(kBlock), 0, st, k, p0, p1, n, shift, hist, nb, k2, a2, b2);
    HIPCHK(hipGetLastError());
    std::swap(k, k2);
    std::swap(p0, a2);
    std::swap(p1, b2);
  }
  if (sync) HIPCHK(hipStreamSynchronize(st));
  dfree(k2);
  dfree(a2);
  dfree(b2);
  dfree(hist);
  dfree(tt);
  dfree(dTot);
}

inline int bitsFor(uint64_t maxValue) {
  int b = 0;
  while (b < 32 && (maxValue >> b)) ++b;
  return std::max(8, (b + 7) / 8 * 8);
}


// Domain and subdivision scheme of a tree: ImplicitBoxTree.hpp:27-45 (default scheme),
// :3512-3552 setSdScheme (periodic extension), :574-587 adaptSdCount. False if sd has values >= dim.
inline bool buildDomain(int dim, const double* center, const double* radius, const uint8_t* sd, int nsd, Domain& D) {
  std::vector<uint8_t> scheme;
  if (sd) {
    for (int i = 0; i < nsd; ++i)
      if (sd[i] >= dim) return false;
    scheme.resize(kMaxDepth + 1);
    for (int i = 0; i < kMaxDepth + 1; ++i) scheme[i] = nsd > 0 ? sd[i % nsd] : uint8_t(i % dim);
  } else {
    scheme.resize(kMaxDepth);
    for (int i = 0; i < kMaxDepth; ++i) scheme[i] = uint8_t(i % dim);
  }
  std::memset(&D, 0, sizeof(D));
  for (int j = 0; j < dim; ++j) {
    D.center[j] = center[j];
    D.radius[j] = radius[j];
  }
  D.isDefault = 1;
  for (size_t i = 0; i < scheme.size(); ++i) {
    D.sdScheme[i] = scheme[i];
    if (scheme[i] != uint8_t(uint8_t(i) % dim)) D.isDefault = 0;
    for (int j = 0; j < dim; ++j) D.sdCount[(i + 1) * dim + j] = D.sdCount[i * dim + j];
    ++D.sdCount[(i + 1) * dim + scheme[i]];
  }
  return true;
}

// Map parameters of a map kind; missing trailing parameters take B12's defaults
// (Map_Henon.h:32, Map_Lorenz.h:77-78). For the forced Duffing map the state-independent time
// loop of MapTools.hpp:54-56 is evaluated here into `tab` (the caller uploads it).
inline void buildMapParams(int dim, int mk, const double* mp, int nmp, MapParams& M, std::vector<ForcingStep>& tab) {
  auto P = [&](int i, double d) { return i < nmp ? mp[i] : d; };
  M = MapParams{};
  M.kind = mk;
      switch (mk) {
        case MAP_IDENTITY: break;
        case MAP_HENON:
          if (dim != 2) throw HipFail(B12X_E_ARG, "Henon needs dim 2");
          M.p[0] = P(0, 1.4); M.p[1] = P(1, 0.3); M.steps = uint32_t(P(2, 1));
          break;
        case MAP_LORENZ: {
          if (dim != 3) throw HipFail(B12X_E_ARG, "Lorenz needs dim 3");
          const double h = P(3, 0.01);
          M.p[0] = P(0, 10.0); M.p[1] = P(1, 28.0); M.p[2] = P(2, 8.0 / 3.0);
          M.p[3] = h; M.p[4] = h / 6.0; M.p[5] = h / 2.0; M.p[6] = h / 3.0;
          M.steps = uint32_t(P(4, 20));
          break;
        }
        case MAP_HENON_HEILES: {
          if (dim != 4) throw HipFail(B12X_E_ARG, "Henon-Heiles needs dim 4");
          M.order = int(P(0, 2));
          if (M.order < 1 || M.order > 4) throw HipFail(B12X_E_ARG, "symplectic order must be 1..4");
          M.p[0] = P(1, 0.02); M.steps = uint32_t(P(2, 50));
          M.p[2] = 0x1.428a2f98d728bp+0;  // cbrt(2) correctly rounded (MapTools.hpp:468)
          M.p[1] = 1.0 / (2.0 - M.p[2]);
          break;
        }
        case MAP_DUFFING: {
          if (dim != 2) throw HipFail(B12X_E_ARG, "Duffing needs dim 2");
          M.p[0] = P(0, 0.25); M.p[1] = P(1, -1.0); M.p[2] = P(2, 1.0);
          const double gamma = P(3, 0.3), omega = P(4, 1.0), t0 = P(5, 0.0), t1 = P(6, 6.283185307179586),
                       hIn = P(7, 6.283185307179586 / 100.0);
          if (!(hIn > 0.0)) throw HipFail(B12X_E_ARG, "Duffing step must be positive");
          // MapTools.hpp:54-56: the state-independent time loop, evaluated once on the host
          for (double tt = t0; tt < t1; tt += hIn) {
            const double h = (t1 - tt) < hIn ? (t1 - tt) : hIn;
            ForcingStep st;
            st.h = h;
            st.h6 = h / 6.0;
            st.h2 = h / 2.0;
            st.h3 = h / 3.0;
            st.f0 = gamma * std::cos(omega * tt);
            st.fh = gamma * std::cos(omega * (tt + h / 2.0));
            st.f1 = gamma * std::cos(omega * (tt + h));
            tab.push_back(st);
            if (tab.size() > (1u << 24)) throw HipFail(B12X_E_ARG, "Duffing time loop too long");
          }
          M.nForcing = uint32_t(tab.size());
          break;
        }
        default: throw HipFail(B12X_E_UNSUPPORTED, "map kind");
      }
}

// Grid parameters (Grids.h:333-370). InnerGrid offsets depend on i mod k only, so the k offsets
// are evaluated here with Grids.hpp:112-117's expression into `table` (the caller uploads it).
inline void buildGrid(int dim, int gk, const uint64_t* gp, int ngp, GridParams& G, std::vector<double>& table) {
  G = GridParams{};
  G.kind = gk;
  if (gk == GRID_INNER) {
    const uint64_t n = ngp > 0 ? gp[0] : 1;
    uint64_t k = 1;
    for (int j = 0; j < dim; ++j) k *= n;
    if (n == 0 || k > (1u << 20)) throw HipFail(B12X_E_ARG, "InnerGrid size");
    G.k = uint32_t(k);
    table.resize(k * dim);
    for (uint64_t i = 0; i < k; ++i) {
      uint64_t npd = 1;
      for (int j = 0; j < dim; ++j) {
        table[i * dim + j] = (double((i / npd) % n) * 2.0 + 1.0 - double(n)) / double(n);
        npd *= n;
      }
    }
  } else if (gk == GRID_UNIFORM) {
    G.k = uint32_t(ngp > 0 ? gp[0] : 1);
    G.seed = ngp > 1 ? gp[1] : 0;
    if (G.k == 0) throw HipFail(B12X_E_ARG, "UniformGrid size");
  } else {
    throw HipFail(B12X_E_UNSUPPORTED, "grid kind");
  }
}

}  // namespace b12x
