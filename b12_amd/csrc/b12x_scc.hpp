// Strongly connected components of the directed graph of a sparse matrix (an edge u -> v for
// every entry (u, v)), as the chain recurrent set step needs them (chainRecurrentSet.m:18-44,
// where MATLAB's graphconncomp labels the components of the transition graph).
//
// The graph is held as in-lists: the sources of the edges into v are src[inPtr[v] ..
// inPtr[v+1]), which is the column-major order of the matrix (the transition matrix's own
// order). Every phase is a sequence of vertex-parallel passes, HBM-bound: a pass streams inPtr
// and the in-lists (4 bytes per vertex and per edge) and gathers one 4-byte colour per edge at
// random, which stays in the L2 and the Infinity Cache for the vertex counts of box covers.
// Repeated until nothing is left:
//   1. trim: a vertex without an edge from or to another unassigned vertex is a component of
//      its own (self loops do not count);
//   2. colouring: every unassigned vertex starts with its own index as colour and pulls the
//      largest colour of its in-neighbours, so it ends with the largest index of the unassigned
//      vertices that reach it (assigned vertices hold colour 0, which never wins);
//   3. a vertex whose colour is its own index is the largest of its component; the vertices of
//      its colour that reach it backwards are that component, labelled with the root's index:
//      marks spread from the roots to in-neighbours of equal colour, each marked vertex pushing
//      once (frontier state 1, done state 2).
// Each round assigns at least the component of the largest unassigned vertex. A component's
// label is therefore the largest vertex index in it, which makes the labelling canonical.
// Passes are launched from the host in groups and no kernel waits on another block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "b12x_host.hpp"
#include "b12x_kernels.hpp"

namespace b12x {

constexpr uint32_t kSccNone = 0xFFFFFFFFu;  // vertex not yet assigned to a component

// trim, part 1: for unassigned v, whether an unassigned u != v has an edge into v (hasIn[v]);
// each such u has an edge out (hasOut[u])
static __global__ __launch_bounds__(kBlock) void sccDegreeKernel(const uint32_t* inPtr, const uint32_t* src,
                                                                uint32_t n, const uint32_t* lab, uint8_t* hasOut,
                                                                uint8_t* hasIn) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n || lab[v] != kSccNone) return;
  bool any = false;
  for (uint32_t e = inPtr[v], e1 = inPtr[v + 1]; e < e1; ++e) {
    const uint32_t u = src[e];
    if (u >= n || u == v || lab[u] != kSccNone) continue;
    any = true;
    hasOut[u] = 1;
  }
  if (any) hasIn[v] = 1;
}

// trim, part 2: an unassigned vertex lacking either kind of edge is a component by itself (its
// colour becomes 0); the marks are cleared for the next pass. *removed (and *removedHere when
// set) count the vertices assigned
static __global__ __launch_bounds__(kBlock) void sccTrimKernel(uint32_t n, uint32_t* lab, uint32_t* col,
                                                              uint8_t* hasOut, uint8_t* hasIn, uint32_t* removed,
                                                              uint32_t* removedHere) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (v < n && lab[v] == kSccNone) {
    if (!hasOut[v] || !hasIn[v]) {
      lab[v] = v;
      col[v] = 0;
      c = 1;
    }
    hasOut[v] = 0;
    hasIn[v] = 0;
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) {
    atomicAdd(removed, c);
    if (removedHere) atomicAdd(removedHere, c);
  }
}

// colouring start: colour = own index, no marks
static __global__ __launch_bounds__(kBlock) void sccColorInitKernel(uint32_t n, const uint32_t* lab, uint32_t* col,
                                                                   uint8_t* vis) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n || lab[v] != kSccNone) return;
  col[v] = v;
  vis[v] = 0;
}

// one colouring pass: an unassigned v takes the largest colour of its in-neighbours when it
// exceeds its own; *changed (when set) is set when a colour grew
static __global__ __launch_bounds__(kBlock) void sccColorKernel(const uint32_t* inPtr, const uint32_t* src, uint32_t n,
                                                               const uint32_t* lab, uint32_t* col,
                                                               uint32_t* changed) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  bool grew = false;
  if (v < n && lab[v] == kSccNone) {
    const uint32_t c = col[v];
    uint32_t m = c;
    uint32_t e = inPtr[v];
    const uint32_t e1 = inPtr[v + 1];
    for (; e + 4 <= e1; e += 4) {  // four independent gathers in flight per lane
      const uint32_t u0 = src[e], u1 = src[e + 1], u2 = src[e + 2], u3 = src[e + 3];
      const uint32_t c0 = u0 < n ? col[u0] : 0u, c1 = u1 < n ? col[u1] : 0u;
      const uint32_t c2 = u2 < n ? col[u2] : 0u, c3 = u3 < n ? col[u3] : 0u;
      m = max(m, max(max(c0, c1), max(c2, c3)));
    }
    for (; e < e1; ++e) {
      const uint32_t u = src[e];
      if (u < n) m = max(m, col[u]);
    }
    if (m > c) {
      col[v] = m;
      grew = true;
    }
  }
  if (changed && __any(grew) && (threadIdx.x & (kWave - 1)) == 0 &&
      __hip_atomic_load(changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
    atomicOr(changed, 1u);  // none once the flag reads as set (same-address atomics serialise)
}

// the roots: unassigned vertices whose colour is their own index (frontier state 1)
static __global__ __launch_bounds__(kBlock) void sccRootKernel(uint32_t n, const uint32_t* lab, const uint32_t* col,
                                                              uint8_t* vis) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n || lab[v] != kSccNone) return;
  vis[v] = col[v] == v ? 1 : 0;
}

// one backward pass: every frontier vertex v marks its unmarked in-neighbours of its colour and
// leaves the frontier; *changed (when set) is set when a mark was added
static __global__ __launch_bounds__(kBlock) void sccBackKernel(const uint32_t* inPtr, const uint32_t* src, uint32_t n,
                                                              const uint32_t* lab, const uint32_t* col, uint8_t* vis,
                                                              uint32_t* changed) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  bool grew = false;
  if (v < n && vis[v] == 1 && lab[v] == kSccNone) {
    vis[v] = 2;
    const uint32_t c = col[v];
    for (uint32_t e = inPtr[v], e1 = inPtr[v + 1]; e < e1; ++e) {
      const uint32_t u = src[e];
      if (u >= n || vis[u] || col[u] != c || lab[u] != kSccNone) continue;
      vis[u] = 1;
      grew = true;
    }
  }
  if (changed && __any(grew) && (threadIdx.x & (kWave - 1)) == 0 &&
      __hip_atomic_load(changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
    atomicOr(changed, 1u);  // none once the flag reads as set (same-address atomics serialise)
}

// the marked vertices take their colour as label (and colour 0); *left counts the vertices
// still unassigned
static __global__ __launch_bounds__(kBlock) void sccAssignKernel(uint32_t n, uint32_t* lab, uint32_t* col,
                                                                const uint8_t* vis, uint32_t* left) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (v < n && lab[v] == kSccNone) {
    if (vis[v]) {
      lab[v] = col[v];
      col[v] = 0;
    } else {
      c = 1;
    }
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) atomicAdd(left, c);
}

// components of more than one vertex: a vertex whose label is not its own index marks its
// component's root (the label) as shared; the mark is written once, by the first lane that finds
// it unset, so a giant component does not serialise on one address
static __global__ __launch_bounds__(kBlock) void sccShareKernel(uint32_t n, const uint32_t* lab, uint8_t* shared) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n) return;
  const uint32_t l = lab[v];
  if (l < n && l != v && !shared[l]) shared[l] = 1;
}

// per vertex whether its component has more than one vertex; counts of the components (their
// roots) and of the shared ones
static __global__ __launch_bounds__(kBlock) void sccNontrivialKernel(uint32_t n, const uint32_t* lab,
                                                                    const uint8_t* shared, uint8_t* keep,
                                                                    uint32_t* nComp, uint32_t* nBig) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0, b = 0;
  if (v < n) {
    const uint32_t l = lab[v];
    if (keep) keep[v] = (l != v || shared[v]) ? 1 : 0;
    if (l == v) {
      c = 1;
      b = shared[v] ? 1u : 0u;
    }
  }
  c = blockReduceSum(c);
  b = blockReduceSum(b);
  if (threadIdx.x == 0) {
    if (c) atomicAdd(nComp, c);
    if (b) atomicAdd(nBig, b);
  }
}

// Host side: two counters in device memory, read back together.
struct SccCounters {
  uint32_t* d = nullptr;
  explicit SccCounters() { d = dalloc<uint32_t>(2); }
  ~SccCounters() { dfree(d); }
  void read(hipStream_t st, uint32_t h[2]) const {
    HIPCHK(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  void clear(hipStream_t st) const { HIPCHK(hipMemsetAsync(d, 0, 8, st)); }
};

struct SccStats {
  uint32_t components = 0, nontrivial = 0, rounds = 0, passes = 0, checks = 0;
};

constexpr int kSccMaxGroup = 16;  // passes launched between two reads of the change counters

// Labels lab[0..n) (device) of the strongly connected components of the graph whose edges into
// v come from src[inPtr[v] .. inPtr[v+1]) (device; sources >= n are ignored). Each label is the
// largest vertex index of its component. keep (device, may be null) receives per vertex whether
// its component holds more than one vertex.
// Each phase launches its passes in groups (1, 2, 4, ... up to kSccMaxGroup) and reads the change
// counter of the group's last pass only: a last pass that changed nothing proves the fixed point,
// at the price of at most a group's worth of passes beyond it.
inline SccStats sccLabels(hipStream_t st, uint32_t n, const uint32_t* inPtr, const uint32_t* src, uint32_t* lab,
                          uint8_t* keep) {
  SccStats s;
  if (n == 0) return s;
  const uint32_t vb = ceilDiv(n, kBlock);
  DevScope ds;
  uint32_t* col = ds.alloc<uint32_t>(n);
  uint8_t* hasOut = ds.alloc<uint8_t>(n);
  uint8_t* hasIn = ds.alloc<uint8_t>(n);
  uint8_t* vis = ds.alloc<uint8_t>(n);
  SccCounters cnt;
  uint32_t h[2];
  HIPCHK(hipMemsetAsync(lab, 0xFF, size_t(n) * 4, st));
  HIPCHK(hipMemsetAsync(col, 0, size_t(n) * 4, st));
  HIPCHK(hipMemsetAsync(hasOut, 0, n, st));
  HIPCHK(hipMemsetAsync(hasIn, 0, n, st));
  // a colour or a mark changes at most n times along a path, a trim pass removes a vertex or ends
  // the trimming: no phase needs more than n + 1 passes
  const uint64_t passCap = uint64_t(n) + kSccMaxGroup + 2;
  auto grow = [](int g) { return std::min(2 * g, kSccMaxGroup); };
  uint32_t left = n;
  for (;;) {
    // trim while it removes vertices (each pass peels the sources and sinks of what is left)
    uint64_t phase = 0;
    for (int g = 1; left; g = grow(g)) {
      cnt.clear(st);
      for (int j = 0; j < g; ++j) {
        hipLaunchKernelGGL(sccDegreeKernel, dim3(vb), dim3(kBlock), 0, st, inPtr, src, n, (const uint32_t*)lab,
                           hasOut, hasIn);
        hipLaunchKernelGGL(sccTrimKernel, dim3(vb), dim3(kBlock), 0, st, n, lab, col, hasOut, hasIn, cnt.d,
                           j == g - 1 ? cnt.d + 1 : nullptr);
      }
      HIPCHK(hipGetLastError());
      cnt.read(st, h);
      s.passes += g;
      ++s.checks;
      left -= h[0];
      if (h[1] == 0) break;
      if ((phase += g) > passCap) throw HipFail(B12X_E_HIP, "strong components: trimming did not settle");
    }
    if (left == 0) break;
    ++s.rounds;
    hipLaunchKernelGGL(sccColorInitKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, col, vis);
    phase = 0;
    for (int g = 1;; g = grow(g)) {
      for (int j = 0; j < g; ++j) {
        if (j == g - 1) cnt.clear(st);
        hipLaunchKernelGGL(sccColorKernel, dim3(vb), dim3(kBlock), 0, st, inPtr, src, n, (const uint32_t*)lab, col,
                           j == g - 1 ? cnt.d : nullptr);
      }
      HIPCHK(hipGetLastError());
      cnt.read(st, h);
      s.passes += g;
      ++s.checks;
      if (!h[0]) break;
      if ((phase += g) > passCap) throw HipFail(B12X_E_HIP, "strong components: colouring did not settle");
    }
    hipLaunchKernelGGL(sccRootKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, (const uint32_t*)col,
                       vis);
    phase = 0;
    for (int g = 1;; g = grow(g)) {
      for (int j = 0; j < g; ++j) {
        if (j == g - 1) cnt.clear(st);
        hipLaunchKernelGGL(sccBackKernel, dim3(vb), dim3(kBlock), 0, st, inPtr, src, n, (const uint32_t*)lab,
                           (const uint32_t*)col, vis, j == g - 1 ? cnt.d : nullptr);
      }
      HIPCHK(hipGetLastError());
      cnt.read(st, h);
      s.passes += g;
      ++s.checks;
      if (!h[0]) break;
      if ((phase += g) > passCap) throw HipFail(B12X_E_HIP, "strong components: backward marking did not settle");
    }
    cnt.clear(st);
    hipLaunchKernelGGL(sccAssignKernel, dim3(vb), dim3(kBlock), 0, st, n, lab, col, (const uint8_t*)vis, cnt.d);
    HIPCHK(hipGetLastError());
    cnt.read(st, h);
    ++s.checks;
    if (h[0] >= left) throw HipFail(B12X_E_HIP, "strong components: a round assigned no vertex");
    left = h[0];
    if (left == 0) break;
  }
  // the shared marks (hasOut reused), the nontrivial marks and the counts
  HIPCHK(hipMemsetAsync(hasOut, 0, n, st));
  cnt.clear(st);
  hipLaunchKernelGGL(sccShareKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, hasOut);
  hipLaunchKernelGGL(sccNontrivialKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab,
                     (const uint8_t*)hasOut, keep, cnt.d, cnt.d + 1);
  HIPCHK(hipGetLastError());
  cnt.read(st, h);
  s.components = h[0];
  s.nontrivial = h[1];
  return s;
}

}  // namespace b12x
