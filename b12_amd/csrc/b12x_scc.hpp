// Strongly connected components of the directed graph of a sparse matrix (an edge u -> v for
// every entry (u, v)), as the chain recurrent set step needs them (chainRecurrentSet.m:18-44,
// where MATLAB's graphconncomp labels the components of the transition graph).
//
// The graph is an edge list in HBM (the COO arrays, 8 bytes per edge) and every phase is a
// sequence of edge-parallel passes over it, HBM-bound: the per-vertex arrays (labels, colours,
// marks: 6 bytes per vertex) are gathered at random but stay in the L2 and the Infinity Cache
// for the vertex counts of box covers. Repeated until nothing changes:
//   1. trim: a vertex without an edge from or to another unassigned vertex is a component of
//      its own (self loops do not count);
//   2. colouring: every unassigned vertex starts with its own index as colour and colours are
//      pushed along the edges with atomicMax, so a vertex ends with the largest index of the
//      unassigned vertices that reach it;
//   3. a vertex whose colour is its own index is the largest of its component; the vertices of
//      its colour that reach it backwards are that component, labelled with the root's index.
// Each round assigns at least the component of the largest unassigned vertex. A component's
// label is therefore the largest vertex index in it, which makes the labelling canonical.
// Passes are launched from the host, which reads one change flag after each: no kernel waits
// on another block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "b12x_host.hpp"
#include "b12x_kernels.hpp"

namespace b12x {

constexpr uint32_t kSccNone = 0xFFFFFFFFu;  // vertex not yet assigned to a component
constexpr uint32_t kSccEdgeBlocks = 8192;   // grid of the edge passes (grid-stride loops)

// trim, part 1: mark the unassigned vertices with an edge to (out) or from (in) another one
static __global__ __launch_bounds__(kBlock) void sccDegreeKernel(const uint32_t* src, const uint32_t* dst, uint64_t m,
                                                                uint32_t n, const uint32_t* lab, uint8_t* hasOut,
                                                                uint8_t* hasIn) {
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < m; e += uint64_t(gridDim.x) * kBlock) {
    const uint32_t u = src[e], v = dst[e];
    if (u >= n || v >= n || u == v) continue;
    if (lab[u] != kSccNone || lab[v] != kSccNone) continue;
    hasOut[u] = 1;
    hasIn[v] = 1;
  }
}

// trim, part 2: an unassigned vertex lacking either kind of edge is a component by itself; the
// marks are cleared for the next pass, *removed counts the vertices assigned
static __global__ __launch_bounds__(kBlock) void sccTrimKernel(uint32_t n, uint32_t* lab, uint8_t* hasOut,
                                                              uint8_t* hasIn, uint32_t* removed) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (v < n && lab[v] == kSccNone) {
    if (!hasOut[v] || !hasIn[v]) {
      lab[v] = v;
      c = 1;
    }
    hasOut[v] = 0;
    hasIn[v] = 0;
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) atomicAdd(removed, c);
}

// colouring start: colour = own index, no marks
static __global__ __launch_bounds__(kBlock) void sccColorInitKernel(uint32_t n, const uint32_t* lab, uint32_t* col,
                                                                   uint8_t* vis) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n || lab[v] != kSccNone) return;
  col[v] = v;
  vis[v] = 0;
}

// one colouring pass: col[v] = max(col[v], col[u]) over the edges u -> v between unassigned
// vertices; *changed is set when a colour grew
static __global__ __launch_bounds__(kBlock) void sccColorKernel(const uint32_t* src, const uint32_t* dst, uint64_t m,
                                                               uint32_t n, const uint32_t* lab, uint32_t* col,
                                                               uint32_t* changed) {
  bool grew = false;
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < m; e += uint64_t(gridDim.x) * kBlock) {
    const uint32_t u = src[e], v = dst[e];
    if (u >= n || v >= n || u == v) continue;
    if (lab[u] != kSccNone || lab[v] != kSccNone) continue;
    const uint32_t cu = col[u];
    if (cu > col[v] && atomicMax(col + v, cu) < cu) grew = true;
  }
  if (__any(grew) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(changed, 1u);
}

// the roots: unassigned vertices whose colour is their own index
static __global__ __launch_bounds__(kBlock) void sccRootKernel(uint32_t n, const uint32_t* lab, const uint32_t* col,
                                                              uint8_t* vis) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n || lab[v] != kSccNone) return;
  vis[v] = col[v] == v ? 1 : 0;
}

// one backward pass: u is marked when it has an edge u -> v to a marked vertex of its colour
static __global__ __launch_bounds__(kBlock) void sccBackKernel(const uint32_t* src, const uint32_t* dst, uint64_t m,
                                                              uint32_t n, const uint32_t* lab, const uint32_t* col,
                                                              uint8_t* vis, uint32_t* changed) {
  bool grew = false;
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < m; e += uint64_t(gridDim.x) * kBlock) {
    const uint32_t u = src[e], v = dst[e];
    if (u >= n || v >= n || u == v) continue;
    if (lab[u] != kSccNone || lab[v] != kSccNone) continue;
    if (vis[v] && !vis[u] && col[u] == col[v]) {
      vis[u] = 1;
      grew = true;
    }
  }
  if (__any(grew) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(changed, 1u);
}

// the marked vertices take their colour as label; *left counts the vertices still unassigned
static __global__ __launch_bounds__(kBlock) void sccAssignKernel(uint32_t n, uint32_t* lab, const uint32_t* col,
                                                                const uint8_t* vis, uint32_t* left) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (v < n && lab[v] == kSccNone) {
    if (vis[v]) lab[v] = col[v];
    else c = 1;
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) atomicAdd(left, c);
}

// per label, the number of vertices; then per vertex whether its component has more than one
static __global__ __launch_bounds__(kBlock) void sccSizeKernel(uint32_t n, const uint32_t* lab, uint32_t* size) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n && lab[v] < n) atomicAdd(size + lab[v], 1u);
}

static __global__ __launch_bounds__(kBlock) void sccNontrivialKernel(uint32_t n, const uint32_t* lab,
                                                                    const uint32_t* size, uint8_t* keep,
                                                                    uint32_t* nComp, uint32_t* nBig) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0, b = 0;
  if (v < n && lab[v] < n) {
    const uint32_t s = size[lab[v]];
    if (keep) keep[v] = s > 1 ? 1 : 0;
    if (lab[v] == v) {
      c = 1;
      b = s > 1 ? 1u : 0u;
    }
  }
  c = blockReduceSum(c);
  b = blockReduceSum(b);
  if (threadIdx.x == 0) {
    if (c) atomicAdd(nComp, c);
    if (b) atomicAdd(nBig, b);
  }
}

// Host side: counters in device memory, read back after each pass.
struct SccCounters {
  uint32_t* d = nullptr;  // [0] change flag / counts, [1] second count
  explicit SccCounters() { d = dalloc<uint32_t>(2); }
  ~SccCounters() { dfree(d); }
  uint32_t read(hipStream_t st, int i = 0) const {
    uint32_t h = 0;
    HIPCHK(hipMemcpyAsync(&h, d + i, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
  }
  void clear(hipStream_t st) const { HIPCHK(hipMemsetAsync(d, 0, 8, st)); }
};

struct SccStats {
  uint32_t components = 0, nontrivial = 0, rounds = 0, passes = 0;
};

// Labels lab[0..n) (device) of the strongly connected components of the graph with the m edges
// src[e] -> dst[e] (device); edges naming a vertex >= n are ignored. Each label is the largest
// vertex index of its component. keep (device, may be null) receives per vertex whether its
// component holds more than one vertex.
inline SccStats sccLabels(hipStream_t st, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                          uint32_t* lab, uint8_t* keep) {
  SccStats s;
  if (n == 0) return s;
  const uint32_t vb = ceilDiv(n, kBlock);
  const uint32_t eb = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(ceilDiv(m, kBlock), kSccEdgeBlocks)));
  uint32_t* col = dalloc<uint32_t>(n);
  uint8_t* hasOut = dalloc<uint8_t>(n);
  uint8_t* hasIn = dalloc<uint8_t>(n);
  uint8_t* vis = dalloc<uint8_t>(n);
  SccCounters cnt;
  HIPCHK(hipMemsetAsync(lab, 0xFF, size_t(n) * 4, st));
  HIPCHK(hipMemsetAsync(hasOut, 0, n, st));
  HIPCHK(hipMemsetAsync(hasIn, 0, n, st));
  // every pass is bounded: a colour or a mark can change at most n times along a path
  const uint64_t passCap = uint64_t(n) + 2;
  uint32_t left = n;
  for (;;) {
    // trim while it removes vertices (each pass peels the sources and sinks of what is left)
    for (uint64_t t = 0; left; ++t) {
      if (t > passCap) throw HipFail(B12X_E_HIP, "strong components: trimming did not settle");
      cnt.clear(st);
      if (m) hipLaunchKernelGGL(sccDegreeKernel, dim3(eb), dim3(kBlock), 0, st, src, dst, m, n, (const uint32_t*)lab,
                                hasOut, hasIn);
      hipLaunchKernelGGL(sccTrimKernel, dim3(vb), dim3(kBlock), 0, st, n, lab, hasOut, hasIn, cnt.d);
      HIPCHK(hipGetLastError());
      const uint32_t removed = cnt.read(st);
      ++s.passes;
      left -= removed;
      if (removed == 0) break;
    }
    if (left == 0) break;
    ++s.rounds;
    hipLaunchKernelGGL(sccColorInitKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, col, vis);
    for (uint64_t p = 0;; ++p) {
      if (p > passCap) throw HipFail(B12X_E_HIP, "strong components: colouring did not settle");
      cnt.clear(st);
      hipLaunchKernelGGL(sccColorKernel, dim3(eb), dim3(kBlock), 0, st, src, dst, m, n, (const uint32_t*)lab, col,
                         cnt.d);
      HIPCHK(hipGetLastError());
      ++s.passes;
      if (!cnt.read(st)) break;
    }
    hipLaunchKernelGGL(sccRootKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, (const uint32_t*)col,
                       vis);
    for (uint64_t p = 0;; ++p) {
      if (p > passCap) throw HipFail(B12X_E_HIP, "strong components: backward marking did not settle");
      cnt.clear(st);
      hipLaunchKernelGGL(sccBackKernel, dim3(eb), dim3(kBlock), 0, st, src, dst, m, n, (const uint32_t*)lab,
                         (const uint32_t*)col, vis, cnt.d);
      HIPCHK(hipGetLastError());
      ++s.passes;
      if (!cnt.read(st)) break;
    }
    cnt.clear(st);
    hipLaunchKernelGGL(sccAssignKernel, dim3(vb), dim3(kBlock), 0, st, n, lab, (const uint32_t*)col,
                       (const uint8_t*)vis, cnt.d);
    HIPCHK(hipGetLastError());
    const uint32_t nowLeft = cnt.read(st);
    if (nowLeft >= left) throw HipFail(B12X_E_HIP, "strong components: a round assigned no vertex");
    left = nowLeft;
    if (left == 0) break;
  }
  // sizes, the nontrivial marks and the counts (col reused for the sizes)
  HIPCHK(hipMemsetAsync(col, 0, size_t(n) * 4, st));
  cnt.clear(st);
  hipLaunchKernelGGL(sccSizeKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab, col);
  hipLaunchKernelGGL(sccNontrivialKernel, dim3(vb), dim3(kBlock), 0, st, n, (const uint32_t*)lab,
                     (const uint32_t*)col, keep, cnt.d, cnt.d + 1);
  HIPCHK(hipGetLastError());
  s.components = cnt.read(st, 0);
  s.nontrivial = cnt.read(st, 1);
  dfree(col);
  dfree(hasOut);
  dfree(hasIn);
  dfree(vis);
  return s;
}

}  // namespace b12x
