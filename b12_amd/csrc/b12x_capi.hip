// Host implementation of the b12x C ABI (include/b12x_c_api.h): tree, scheme and matrix handles,
// device buffers, and the launch sequences of every operation. Reference citations are relative
// to B12c/src/include/.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/b12x_c_api.h"
#include "b12x_host.hpp"
#include "b12x_kernels.hpp"
#include "b12x_launch.hpp"
#include "b12x_scc.hpp"

using namespace b12x;


// ------------------------------------------------------------------------------------------------
// Handles
// ------------------------------------------------------------------------------------------------
struct b12x_scheme_s {
  uint64_t id = 0;  // unique per scheme (a captured graph records the scheme it was built for)
  int dim = 0;
  GridParams grid{};
  MapParams map{};
  double* dTable = nullptr;
  ForcingStep* dForcing = nullptr;
  ~b12x_scheme_s() {
    dfree(dTable);
    dfree(dForcing);
  }
};

// tags of pooled graph executables (ObjectPool): a tree's large and small step graphs
constexpr int kGraphKindStep = 0;
// B12X_GRAPH_REUSE=0: every tree instantiates its step graphs anew
inline bool graphReuse() {
  static const bool on = [] {
    const char* e = std::getenv("B12X_GRAPH_REUSE");
    return !(e && e[0] == '0');
  }();
  return on;
}

struct b12x_tree_s {
  int dim = 0;
  int sdEntries = 0;  // entries of the reference's sdScheme_ (64 default, 65 when given)
  Domain dom{};
  hipStream_t stream = nullptr;
  uint32_t n = 0;      // number of leaves
  uint32_t cap = 0;    // capacity of every per-leaf buffer
  uint64_t* key[2] = {nullptr, nullptr};
  uint8_t* depth[2] = {nullptr, nullptr};
  uint8_t* flags[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t* hit = nullptr;      // one byte per leaf, all zero between operations
  uint32_t* list = nullptr;    // source-box indices for flag-filtered point iterators
  uint32_t* totals = nullptr;  // per-tile totals of the scans
  uint32_t* scal = nullptr;    // device scalars
  uint32_t* hscal = nullptr;   // pinned host mirror of scal
  uint64_t* rootKey = nullptr; // the surrounding box as a one-element source
  uint64_t* samples = nullptr; // descent samples of the leaf keys (kDescentSamples)
  uint64_t* srcKey = nullptr;  // sources of a point iterator at a general depth (truncated leaves)
  uint8_t* srcDepth = nullptr;
  uint32_t srcCap = 0;
  bool timing = false;          // HIP-event timing of image launches
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> eventPool;
  double imageMs = 0.0;
  uint64_t imageLaunches = 0, imagePoints = 0;
  std::vector<double> launchMs;          // each timed image launch's duration, in launch order
  std::vector<uint64_t> launchPoints;    // and its test points
  uint8_t* rootDepth = nullptr;

  // the subdivision step (subdivide, images, keep) captured as a graph, in a small form (one-block
  // scans) and a large one, for levels whose subdivided count fits the form's bound; rebuilt when
  // the buffers move or the scheme or flag changes
  struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t scheme = 0, gen = 0;
    uint8_t sd = 0;
    int cur = -1;
    void drop() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
  } step[2];
  uint64_t bufGen = 0;  // incremented whenever the leaf buffers are reallocated
  // incremented whenever the leaves change (not their flags); the cached depth range belongs to
  // rangeGen
  uint64_t leafGen = 1, rangeGen = 0;
  uint8_t minDepthC = 0, maxDepthC = 0;
  uint64_t hostBufferBytes = uint64_t(4) << 30;  // nHostBufferBytes_ (ImplicitBoxTree.hpp:38)
  uint64_t tmPassBytes = 0;  // scratch bound of a transition-matrix pass (0: a third of the free HBM)
  size_t hscalBytes = 0;     // size of the pinned block hscal

  TreeView view(int b) { return {key[b], depth[b], flags[b]}; }
  ~b12x_tree_s() {
    streamRetire(stream);
    // the step graphs' executables go to the pool (the stream has retired: none is running),
    // where the next tree updates them in place
    for (int f = 0; f < 2; ++f) {
      if (step[f].exec && graphReuse()) {
        ObjectPool::get().giveExec(kGraphKindStep + f, step[f].exec);
        step[f].exec = nullptr;
      }
      step[f].drop();
    }
    for (int b = 0; b < 2; ++b) {
      dfree(key[b]);
      dfree(depth[b]);
      dfree(flags[b]);
    }
    dfree(hit);
    dfree(list);
    dfree(totals);
    dfree(scal);
    dfree(rootKey);
    dfree(rootDepth);
    dfree(samples);
    dfree(srcKey);
    dfree(srcDepth);
    if (hscal) ObjectPool::get().givePinned(hscal, hscalBytes);
    for (auto& p : pending) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
    for (auto e : eventPool) (void)hipEventDestroy(e);
    streamRelease(stream);
  }
};

struct b12x_coo_s {
  uint32_t nRows = 0, nCols = 0;
  uint64_t nnz = 0;
  bool rowMajor = false;  // entries sorted by (row, column)
  bool colMajor = false;  // entries sorted by (column, row)
  uint32_t* rows = nullptr;
  uint32_t* cols = nullptr;
  double* vals = nullptr;
  hipStream_t stream = nullptr;
  ~b12x_coo_s() {
    streamRetire(stream);
    dfree(rows);
    dfree(cols);
    dfree(vals);
    streamRelease(stream);
  }
};

namespace {

constexpr int kScal = 32;
// pinned host words: a mirror of the kScal device scalars, then the log of a batch of captured
// subdivision steps (4 counts per step)
constexpr int kStepBatch = 16;
constexpr int kHostScal = kScal + 4 * kStepBatch;
constexpr int kScalStepLog = 25;    // device slot: steps logged in the current batch
constexpr int kScalStepsLeft = 27;  // device slot: subdivides left in a small-form batch
constexpr int kMaxScanTiles = 1 << 24;

void syncScal(b12x_tree_s& T, int count = kScal) {
  HIPCHK(hipMemcpyAsync(T.hscal, T.scal, count * sizeof(uint32_t), hipMemcpyDeviceToHost, T.stream));
  HIPCHK(hipStreamSynchronize(T.stream));
}

void zeroScal(b12x_tree_s& T) { HIPCHK(hipMemsetAsync(T.scal, 0, kScal * sizeof(uint32_t), T.stream)); }

// Grow every per-leaf buffer to hold at least `need` leaves, keeping the current leaves.
void ensureCap(b12x_tree_s& T, uint64_t need) {
  if (need > 0xFFFFFFFEull) throw HipFail(B12X_E_ARG, "number of leaves exceeds NrBoxes range");
  if (need <= T.cap) return;
  uint64_t nc = std::max<uint64_t>(need, std::min<uint64_t>(uint64_t(T.cap) * 2, 0xFFFFFFFEull));
  nc = std::max<uint64_t>(nc, 1024);
  uint64_t* k[2];
  uint8_t* d[2];
  uint8_t* f[2];
  for (int b = 0; b < 2; ++b) {
    k[b] = dalloc<uint64_t>(nc);
    d[b] = dalloc<uint8_t>(nc);
    f[b] = dalloc<uint8_t>(nc);
  }
  if (T.n) {
    HIPCHK(hipMemcpyAsync(k[0], T.key[T.cur], size_t(T.n) * 8, hipMemcpyDeviceToDevice, T.stream));
    HIPCHK(hipMemcpyAsync(d[0], T.depth[T.cur], T.n, hipMemcpyDeviceToDevice, T.stream));
    HIPCHK(hipMemcpyAsync(f[0], T.flags[T.cur], T.n, hipMemcpyDeviceToDevice, T.stream));
  }
  ++T.bufGen;
  uint8_t* h = dalloc<uint8_t>(nc);
  HIPCHK(hipMemsetAsync(h, 0, nc, T.stream));
  uint32_t* l = dalloc<uint32_t>(nc);
  uint32_t* tt = dalloc<uint32_t>(uint64_t(nc) * 2 / kTile + 16);
  HIPCHK(hipStreamSynchronize(T.stream));
  for (int b = 0; b < 2; ++b) {
    dfree(T.key[b]);
    dfree(T.depth[b]);
    dfree(T.flags[b]);
    T.key[b] = k[b];
    T.depth[b] = d[b];
    T.flags[b] = f[b];
  }
  dfree(T.hit);
  dfree(T.list);
  dfree(T.totals);
  T.hit = h;
  T.list = l;
  T.totals = tt;
  T.cur = 0;
  T.cap = uint32_t(nc);
}

// Enqueues a three-phase scan over n elements (or *nDev, with n an upper bound); the total lands
// in T.scal[scalSlot] and the summed Op::stat values in *stat (when set). Does not synchronise.
// With countOnly, only the count and total passes run (the op's outputs may be unset).
template <class Op>
void runScan(b12x_tree_s& T, const Op& op, uint32_t n, uint32_t* totals, int scalSlot, const uint32_t* nDev = nullptr,
             int statSlot = -1, bool countOnly = false) {
  if (n == 0) return;
  const uint32_t nb = ceilDiv(n, kTile);
  if (nb > kMaxScanTiles) throw HipFail(B12X_E_ARG, "scan too large");
  uint32_t* stats = statSlot >= 0 ? totals + nb : nullptr;  // totals holds 2*nb + 16 entries
  hipLaunchKernelGGL(scanCountKernel<Op>, dim3(nb), dim3(kBlock), 0, T.stream, op, n, nDev, totals, stats);
  hipLaunchKernelGGL(scanTotalsKernel, dim3(1), dim3(kBlock), 0, T.stream, totals, nb, nDev, T.scal + scalSlot,
                     (const uint32_t*)stats, statSlot >= 0 ? T.scal + statSlot : nullptr);
  if (!countOnly) hipLaunchKernelGGL(scanScatterKernel<Op>, dim3(nb), dim3(kBlock), 0, T.stream, op, n, nDev, totals);
  HIPCHK(hipGetLastError());
}

// runScan in one block (short arrays); hostOut: where the block copies scal[0..3] at the end
struct IndexOp : OpBase {  // indices i with sel(i) — flag-filtered source lists and searchCompacted
  const uint8_t* flags;
  const uint8_t* hit;
  uint8_t check;
  uint32_t* out;
  __device__ uint32_t count(uint32_t i) const {
    return hit ? (hit[i] ? 1u : 0u) : (areFlagsSet(flags[i], check) ? 1u : 0u);
  }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const { out[pos] = i; }
};

__global__ void fillU8Kernel(uint8_t* p, uint64_t n, uint8_t v) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fillU32Kernel(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int DIM, int MAP>
__global__ void mapPointsKernel(MapParams m, const double* in, double* out, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double x[DIM], y[DIM];
#pragma unroll
  for (int j = 0; j < DIM; ++j) x[j] = in[i * DIM + j];
  applyMap<DIM, MAP>(m, x, y);
#pragma unroll
  for (int j = 0; j < DIM; ++j) out[i * DIM + j] = y[j];
}

// -- reductions over the leaves -------------------------------------------------------------
struct LeafStats {
  uint8_t flagsAnd, flagsOr, maxDepth, minDepth;
};

__global__ void statsInitKernel(uint32_t* s) {  // flags AND, flags OR, max depth, min depth
  s[0] = 0xFFu;
  s[1] = 0u;
  s[2] = 0u;
  s[3] = 255u;
}

// AND and OR of all flags, maximum and minimum depth (depth-0 forms of countIfFlags,
// ImplicitBoxTree.hpp:1105-1114, and depth(), :1264-1271): a grid-stride pass, reduced per wave
// by shuffles and per block through LDS, then one atomic per value and block
// (64 blocks: the per-block atomics on the four words serialise, so few blocks with a longer
// loop of 4-byte loads are faster)
constexpr uint32_t kStatsBlocks = 64;
__global__ __launch_bounds__(kBlock) void leafStatsKernel(const uint8_t* flags, const uint8_t* depth, uint32_t n,
                                                          uint32_t* out) {
  uint32_t fa = 0xFF, fo = 0, mx = 0, mn = 255;
  const uint32_t n4 = n / 4;  // the leaf arrays are 256-byte aligned
  const uint32_t* f4 = reinterpret_cast<const uint32_t*>(flags);
  const uint32_t* d4 = reinterpret_cast<const uint32_t*>(depth);
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n4; i += gridDim.x * kBlock) {
    const uint32_t f = f4[i], d = d4[i];
#pragma unroll
    for (int b = 0; b < 32; b += 8) {
      const uint32_t fb = (f >> b) & 0xFFu, db = (d >> b) & 0xFFu;
      fa &= fb;
      fo |= fb;
      mx = max(mx, db);
      mn = min(mn, db);
    }
  }
  for (uint32_t i = 4 * n4 + blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const uint32_t f = flags[i], d = depth[i];
    fa &= f;
    fo |= f;
    mx = max(mx, d);
    mn = min(mn, d);
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    fa &= __shfl_xor(fa, o, kWave);
    fo |= __shfl_xor(fo, o, kWave);
    mx = max(mx, (uint32_t)__shfl_xor(mx, o, kWave));
    mn = min(mn, (uint32_t)__shfl_xor(mn, o, kWave));
  }
  __shared__ uint32_t w[4][kBlock / kWave];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    w[0][wave] = fa;
    w[1][wave] = fo;
    w[2][wave] = mx;
    w[3][wave] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / kWave; ++k) {
      fa &= w[0][k];
      fo |= w[1][k];
      mx = max(mx, w[2][k]);
      mn = min(mn, w[3][k]);
    }
    atomicAnd(out + 0, fa);
    atomicOr(out + 1, fo);
    atomicMax(out + 2, mx);
    atomicMin(out + 3, mn);
  }
}

// any difference between the first n leaves of two trees sets *diff
__global__ void sameLeavesKernel(TreeView a, TreeView b, uint32_t n, uint32_t* diff) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool d = i < n && (a.key[i] != b.key[i] || a.depth[i] != b.depth[i] || a.flags[i] != b.flags[i]);
  setFlagOncePerWave(d, diff);
}

// the leaves' statistics into scal[8..11] (enqueued only)
void enqueueLeafStats(b12x_tree_s& T) {
  hipLaunchKernelGGL(statsInitKernel, dim3(1), dim3(1), 0, T.stream, T.scal + 8);
  const uint32_t nb = std::max<uint32_t>(1, std::min<uint32_t>(ceilDiv(T.n, 4 * kBlock), kStatsBlocks));
  hipLaunchKernelGGL(leafStatsKernel, dim3(nb), dim3(kBlock), 0, T.stream, T.flags[T.cur], T.depth[T.cur], T.n,
                     T.scal + 8);
  HIPCHK(hipGetLastError());
}

LeafStats leafStatsRead(const b12x_tree_s& T) {
  return {uint8_t(T.hscal[8]), uint8_t(T.hscal[9]), uint8_t(T.hscal[10]), uint8_t(T.hscal[11])};
}

LeafStats leafStats(b12x_tree_s& T) {
  enqueueLeafStats(T);
  syncScal(T, 12);
  return leafStatsRead(T);
}

// Depth classes of the B12 API: 0 = root, 255 = leaves; a depth on which all leaves lie
// behaves as the leaf depth (ImplicitBoxTree.hpp areAllLeavesOnDepth branches). Other depths up
// to the deepest leaf, and the unsubdividable depth, address the boxes at that depth
// (DC_GENERAL); deeper ones address none (DC_BEYOND: counts 0, searches invalid).
enum DepthClass { DC_ROOT, DC_LEAF, DC_GENERAL, DC_BEYOND };

DepthClass classifyDepth(b12x_tree_s& T, uint8_t d) {
  if (d == 0) return DC_ROOT;
  if (d == kLeafDepth) return DC_LEAF;
  if (d == kUnsubDepth) return DC_GENERAL;
  if (T.rangeGen != T.leafGen) {  // the depth range only changes with the leaves
    const LeafStats s = leafStats(T);
    T.minDepthC = s.minDepth;
    T.maxDepthC = s.maxDepth;
    T.rangeGen = T.leafGen;
  }
  if (T.minDepthC == d && T.maxDepthC == d) return DC_LEAF;
  return d > T.maxDepthC ? DC_BEYOND : DC_GENERAL;
}

// Boxes at a general depth (ImplicitBoxTree.hpp:1383-1473): box index, membership and box start
// of every leaf. Owns its device arrays; callers synchronise the tree's stream before it ends.
constexpr int kScalDepthCount = 16, kScalDepthFlags = 17;

struct DepthMap {
  uint32_t* box = nullptr;
  uint8_t* in = nullptr;
  uint8_t* start = nullptr;
  uint32_t count = 0;
  uint8_t d = 0;
  DepthMap() = default;
  DepthMap(const DepthMap&) = delete;
  DepthMap& operator=(const DepthMap&) = delete;
  ~DepthMap() {
    dfree(box);
    dfree(in);
    dfree(start);
  }
};

void buildDepthMap(b12x_tree_s& T, uint8_t d, DepthMap& M) {
  const uint32_t n = T.n;
  M.d = d;
  M.box = dalloc<uint32_t>(n);
  M.in = dalloc<uint8_t>(n);
  M.start = dalloc<uint8_t>(n);
  hipLaunchKernelGGL(depthMapKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.view(T.cur), n, d, M.box,
                     M.start, M.in);
  HIPCHK(hipGetLastError());
  M.count = scanArray(T.stream, M.box, n, T.totals, T.scal + kScalDepthCount, T.hscal + kScalDepthCount);
  hipLaunchKernelGGL(depthBoxKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, M.box,
                     (const uint8_t*)M.start, n);
  HIPCHK(hipGetLastError());
}

// Flag operation op (0 set, 1 unset, 2 change) on the leaves of the boxes at depth M.d selected
// by the device stencil sel (M.count bytes); returns countIfFlags (ImplicitBoxTree.hpp:2803-2881,
// 3334-3411, 912-1058, 1097-1214)
uint32_t depthFlagsOp(b12x_tree_s& T, const DepthMap& M, const uint8_t* sel, int op, uint8_t a, uint8_t b) {
  const uint32_t n = T.n;
  uint32_t* acc = dalloc<uint32_t>(std::max<uint32_t>(M.count, 1));
  hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(std::max<uint32_t>(M.count, 1), kBlock)), dim3(kBlock), 0, T.stream,
                     acc, uint64_t(std::max<uint32_t>(M.count, 1)), op == 1 ? 0u : 0xFFu);
  hipLaunchKernelGGL(depthCombineKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream,
                     (const uint8_t*)T.flags[T.cur], n, (const uint32_t*)M.box, (const uint8_t*)M.in, op == 1 ? 1 : 0,
                     acc);
  hipLaunchKernelGGL(fillU32Kernel, dim3(1), dim3(kBlock), 0, T.stream, T.scal + kScalDepthFlags, uint64_t(1), 0u);
  hipLaunchKernelGGL(depthFlagsKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.flags[T.cur], n,
                     (const uint32_t*)M.box, (const uint8_t*)M.in, (const uint8_t*)M.start, sel, (const uint32_t*)acc,
                     op, a, b, T.scal + kScalDepthFlags);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(T.hscal + kScalDepthFlags, T.scal + kScalDepthFlags, 4, hipMemcpyDeviceToHost, T.stream));
  HIPCHK(hipStreamSynchronize(T.stream));
  dfree(acc);
  return T.hscal[kScalDepthFlags];
}

// the surrounding box (key k, depth 0, no flags) as entry 0 of the given arrays (flags optional)
__global__ void putRootKernel(uint64_t* key, uint8_t* depth, uint8_t* flags, uint64_t k) {
  key[0] = k;
  depth[0] = 0;
  if (flags) flags[0] = 0;
}

void insertRootIfEmpty(b12x_tree_s& T) {  // ImplicitBoxTree.hpp:1959-1970
  if (T.n) return;
  ensureCap(T, 1);
  hipLaunchKernelGGL(putRootKernel, dim3(1), dim3(1), 0, T.stream, T.key[T.cur], T.depth[T.cur], T.flags[T.cur],
                     keyOf(0, 0));
  HIPCHK(hipGetLastError());
  T.n = 1;
  ++T.leafGen;
  // the only leaf lies at depth 0
  T.minDepthC = T.maxDepthC = 0;
  T.rangeGen = T.leafGen;
}

// -- point iterators --------------------------------------------------------------------------
struct Source {
  const uint64_t* key;
  const uint8_t* depth;
  const uint32_t* list;
  uint64_t nBoxes;
};

Source makeSource(b12x_tree_s& T, uint8_t dIt, uint8_t check) {
  if (dIt == 0) {  // ImplicitBoxTree.hpp:2082-2120 and :1483-1492 (root only if all leaves carry check)
    uint64_t nb = 1;
    if (check != 0) nb = areFlagsSet(leafStats(T).flagsAnd, check) ? 1 : 0;
    return {T.rootKey, T.rootDepth, nullptr, nb};
  }
  const DepthClass dc = classifyDepth(T, dIt);
  if (dc == DC_BEYOND) return {T.key[T.cur], T.depth[T.cur], nullptr, 0};
  if (dc == DC_GENERAL) {  // :1528-1601 the first leaf of every box at dIt, truncated to dIt
    DepthMap M;
    buildDepthMap(T, dIt, M);
    const uint32_t n = T.n;
    uint32_t* ok = nullptr;
    if (check != 0) {
      ok = dalloc<uint32_t>(std::max<uint32_t>(M.count, 1));
      hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(std::max<uint32_t>(M.count, 1), kBlock)), dim3(kBlock), 0,
                         T.stream, ok, uint64_t(std::max<uint32_t>(M.count, 1)), 1u);
      hipLaunchKernelGGL(depthCheckKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream,
                         (const uint8_t*)T.flags[T.cur], n, (const uint32_t*)M.box, (const uint8_t*)M.in, check, ok);
    }
    uint32_t* pos = dalloc<uint32_t>(n);
    hipLaunchKernelGGL(depthSourceCountKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream,
                       (const uint8_t*)M.start, (const uint32_t*)M.box, (const uint32_t*)ok, n, pos);
    HIPCHK(hipGetLastError());
    const uint32_t m = scanArray(T.stream, pos, n, T.totals, T.scal + kScalDepthFlags, T.hscal + kScalDepthFlags);
    if (m > T.srcCap) {
      dfree(T.srcKey);
      dfree(T.srcDepth);
      T.srcKey = dalloc<uint64_t>(m);
      T.srcDepth = dalloc<uint8_t>(m);
      T.srcCap = m;
    }
    hipLaunchKernelGGL(depthSourceWriteKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.view(T.cur),
                       n, dIt, (const uint8_t*)M.start, (const uint32_t*)M.box, (const uint32_t*)ok,
                       (const uint32_t*)pos, T.srcKey, T.srcDepth);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(ok);
    dfree(pos);
    return {T.srcKey, T.srcDepth, nullptr, m};
  }
  if (check == 0) return {T.key[T.cur], T.depth[T.cur], nullptr, T.n};
  IndexOp op{};
  op.flags = T.flags[T.cur];
  op.hit = nullptr;
  op.check = check;
  op.out = T.list;
  runScan(T, op, T.n, T.totals, 0);
  syncScal(T, 1);
  return {T.key[T.cur], T.depth[T.cur], T.list, T.hscal[0]};
}

// Writes the descent samples of the tree's current leaves (count from nDev when given) and
// returns them for ImageArgs::samples. Opt-in (B12X_DESCENT_LDS=1): on MI355X the upper levels
// of the all-global search stay in L1/L2 anyway, and the LDS fill per block plus the sampling
// launch per level measured slightly slower on the Lorenz subdivision (null: global descent).
const uint64_t* stageSamples(b12x_tree_s& T, const uint32_t* nDev) {
  const char* e = std::getenv("B12X_DESCENT_LDS");
  if (!(e && e[0] == '1')) return nullptr;
  if (!T.samples) T.samples = dalloc<uint64_t>(kDescentSamples);
  hipLaunchKernelGGL(sampleKeysKernel, dim3(kDescentSamples / kBlock), dim3(kBlock), 0, T.stream,
                     (const uint64_t*)T.key[T.cur], T.n, nDev, T.samples);
  HIPCHK(hipGetLastError());
  return T.samples;
}

// Image-kernel variants: test points per lane x block size; sameBox: a lane's points are
// consecutive points of one box (used when k is a multiple of ppl, else points bs apart).
struct ImageVariant {
  int ppl, bs;
  bool sameBox;
};
constexpr ImageVariant kImageVariants[] = {{1, 256, false}, {1, 128, false}, {1, 64, false}, {2, 256, false},
                                           {2, 128, false}, {2, 64, false},  {4, 256, false}, {4, 128, false},
                                           {4, 64, false},  {2, 256, true},  {2, 128, true},  {3, 256, true},
                                           {3, 128, true},  {4, 256, true},  {4, 128, true}};
constexpr int kNumImageVariants = int(sizeof(kImageVariants) / sizeof(kImageVariants[0]));
constexpr uint64_t kTuneSamplePoints = 1ull << 21;

template <int DIM, int MAP, int PPL, bool SAMEBOX = false>
void launchImage3(const ImageArgs& a, int gridKind, int bs, hipStream_t st, uint32_t maxBlocks = 1u << 22) {
  const uint64_t per = uint64_t(bs) * PPL;
  const uint32_t blocks = uint32_t(std::min<uint64_t>((a.nPoints + per - 1) / per, maxBlocks));
  if (blocks == 0) return;
  if (gridKind == GRID_INNER)
    hipLaunchKernelGGL((imageKernel<DIM, MAP, GRID_INNER, PPL, SAMEBOX>), dim3(blocks), dim3(bs), 0, st, a);
  else
    hipLaunchKernelGGL((imageKernel<DIM, MAP, GRID_UNIFORM, PPL, SAMEBOX>), dim3(blocks), dim3(bs), 0, st, a);
}

template <int DIM, int MAP>
void launchVariant(const ImageArgs& a, int gridKind, ImageVariant v, hipStream_t st) {
  if (v.sameBox) {
    switch (v.ppl) {
      case 4: launchImage3<DIM, MAP, 4, true>(a, gridKind, v.bs, st); break;
      case 3: launchImage3<DIM, MAP, 3, true>(a, gridKind, v.bs, st); break;
      default: launchImage3<DIM, MAP, 2, true>(a, gridKind, v.bs, st); break;
    }
    return;
  }
  switch (v.ppl) {
    case 4: launchImage3<DIM, MAP, 4>(a, gridKind, v.bs, st); break;
    case 2: launchImage3<DIM, MAP, 2>(a, gridKind, v.bs, st); break;
    default: launchImage3<DIM, MAP, 1>(a, gridKind, v.bs, st); break;
  }
}

// Autotuner: for each (dimension, map kind, grid kind, log2 of the launch's point count) the
// first launch runs every variant once untimed (code loading) and once timed with HIP events on
// a prefix of at most kTuneSamplePoints of its points, and later launches use the fastest.
// Every variant computes the same images and a repeated launch stores the same hit bytes or
// targets, so tuning never changes a result.
// B12X_IMAGE_TUNE=0: one point per lane, 256 threads; B12X_IMAGE_PPL / B12X_IMAGE_BLOCK /
// B12X_IMAGE_SAMEBOX fix a variant. The staged descent and the search among boxes at a general
// depth always take one point per lane.
std::mutex g_tuneMutex;
std::unordered_map<uint64_t, int> g_tuned;

int fixedImageVariant(const ImageArgs& a) {
  if (a.samples || a.depthBox || a.outKey) return 0;
  const char* t = std::getenv("B12X_IMAGE_TUNE");
  const char* p = std::getenv("B12X_IMAGE_PPL");
  const char* b = std::getenv("B12X_IMAGE_BLOCK");
  if (!p && !b && !(t && t[0] == '0')) return -1;
  const char* sb = std::getenv("B12X_IMAGE_SAMEBOX");
  const int ppl = p ? std::atoi(p) : 1, bs = b ? std::atoi(b) : 256;
  const bool same = sb && sb[0] == '1';
  for (int v = 0; v < kNumImageVariants; ++v)
    if (kImageVariants[v].ppl == ppl && kImageVariants[v].bs == bs && kImageVariants[v].sameBox == same) return v;
  return 0;
}

template <int DIM, int MAP>
void launchImage2(const ImageArgs& a, int gridKind, hipStream_t st) {
  if (a.nPoints == 0) return;
  const int fixed = fixedImageVariant(a);
  if (fixed >= 0) return launchVariant<DIM, MAP>(a, gridKind, kImageVariants[fixed], st);
  int bucket = 0;
  while ((a.nPoints >> (bucket + 1)) != 0) ++bucket;
  const uint64_t key = (uint64_t(DIM) << 48) | (uint64_t(MAP) << 40) | (uint64_t(gridKind) << 32) | uint64_t(bucket);
  std::lock_guard<std::mutex> lock(g_tuneMutex);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) return launchVariant<DIM, MAP>(a, gridKind, kImageVariants[it->second], st);
  ImageArgs sample = a;
  sample.nPoints = std::min<uint64_t>(a.nPoints, kTuneSamplePoints);
  for (int v = 0; v < kNumImageVariants; ++v) launchVariant<DIM, MAP>(sample, gridKind, kImageVariants[v], st);
  hipEvent_t ev[kNumImageVariants + 1];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], st));
  for (int v = 0; v < kNumImageVariants; ++v) {
    launchVariant<DIM, MAP>(sample, gridKind, kImageVariants[v], st);
    HIPCHK(hipEventRecord(ev[v + 1], st));
  }
  HIPCHK(hipEventSynchronize(ev[kNumImageVariants]));
  int best = 0;
  float bestMs = 0.0f;
  for (int v = 0; v < kNumImageVariants; ++v) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev[v], ev[v + 1]));
    if (v == 0 || ms < bestMs) {
      best = v;
      bestMs = ms;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  g_tuned[key] = best;
  launchVariant<DIM, MAP>(a, gridKind, kImageVariants[best], st);  // the whole launch
}

// one point per lane, 256 threads, at most maxBlocks blocks (the kernel strides over the rest):
// the form a captured graph holds, where the grid is sized for a bound and the count is read on
// the device
template <int DIM, int MAP>
void launchImageFixed2(const ImageArgs& a, int gridKind, hipStream_t st, uint32_t maxBlocks) {
  if (a.nPoints == 0) return;
  launchImage3<DIM, MAP, 1>(a, gridKind, kImageVariants[0].bs, st, maxBlocks);
}

void launchImageFixed(int dim, const ImageArgs& a, int gridKind, hipStream_t st, uint32_t maxBlocks) {
  switch (a.map.kind) {
    case MAP_IDENTITY:
      if (dim == 1) launchImageFixed2<1, MAP_IDENTITY>(a, gridKind, st, maxBlocks);
      else if (dim == 2) launchImageFixed2<2, MAP_IDENTITY>(a, gridKind, st, maxBlocks);
      else if (dim == 3) launchImageFixed2<3, MAP_IDENTITY>(a, gridKind, st, maxBlocks);
      else launchImageFixed2<4, MAP_IDENTITY>(a, gridKind, st, maxBlocks);
      break;
    case MAP_HENON: launchImageFixed2<2, MAP_HENON>(a, gridKind, st, maxBlocks); break;
    case MAP_LORENZ: launchImageFixed2<3, MAP_LORENZ>(a, gridKind, st, maxBlocks); break;
    case MAP_HENON_HEILES: launchImageFixed2<4, MAP_HENON_HEILES>(a, gridKind, st, maxBlocks); break;
    case MAP_DUFFING: launchImageFixed2<2, MAP_DUFFING>(a, gridKind, st, maxBlocks); break;
    default: throw HipFail(B12X_E_UNSUPPORTED, "map kind");
  }
  HIPCHK(hipGetLastError());
}

void launchImage(int dim, const ImageArgs& a, int gridKind, hipStream_t st) {
  switch (a.map.kind) {
    case MAP_IDENTITY:
      if (dim == 1) launchImage2<1, MAP_IDENTITY>(a, gridKind, st);
      else if (dim == 2) launchImage2<2, MAP_IDENTITY>(a, gridKind, st);
      else if (dim == 3) launchImage2<3, MAP_IDENTITY>(a, gridKind, st);
      else launchImage2<4, MAP_IDENTITY>(a, gridKind, st);
      break;
    case MAP_HENON: launchImage2<2, MAP_HENON>(a, gridKind, st); break;
    case MAP_LORENZ: launchImage2<3, MAP_LORENZ>(a, gridKind, st); break;
    case MAP_HENON_HEILES: launchImage2<4, MAP_HENON_HEILES>(a, gridKind, st); break;
    case MAP_DUFFING: launchImage2<2, MAP_DUFFING>(a, gridKind, st); break;
    default: throw HipFail(B12X_E_UNSUPPORTED, "map kind");
  }
  HIPCHK(hipGetLastError());
}

// The subdivision step of the subdivision algorithm: subdivide(sd) (ImplicitBoxTree.hpp:2883-2930),
// setFlagsByIterators(OverLeaves(grid, map), HIT, -1) (the images of every test point), then
// remove(HIT) and unsetFlags(HIT, -1) as one stable compaction. boundN bounds the leaves before
// the step; countDev holds their count on the device (nullptr: boundN is the count).
// The captured forms hold a fixed image variant with a capped grid (a tree that times its images
// launches eagerly); the small form keeps and subdivides in one block, in place, and logs the
// step's counts there; the large form scans in tiles and logs them with a one-thread kernel.
enum StepForm { STEP_EAGER, STEP_LARGE, STEP_SMALL };
constexpr uint32_t kStepGraphLeaves = 1u << 17;  // bound of the large form's subdivided count
constexpr uint32_t kStepSmallLeaves = 1u << 14;  // bound of the small form's subdivided count
constexpr uint32_t kStepGraphBlocks = 8192;      // image blocks of a captured step (they stride)
constexpr int kScalStepIn = 20;                  // pinned slot staging the leaf count

// Batched steps are launched on the stream (back to back, the device idle only while a kernel
// drains); B12X_STEP_LAUNCH=graph replays a captured step instead, which measured slower: the
// device idles about 8 µs between two graph launches.
bool stepLaunchStream() {
  const char* e = std::getenv("B12X_STEP_LAUNCH");
  return !(e && std::strcmp(e, "graph") == 0);
}

bool stepGraphsOn(const b12x_tree_s& T) {
  if (T.timing) return false;  // image timing brackets each eager image launch with events
  const char* ge = std::getenv("B12X_GRAPHS");
  if (ge && ge[0] == '0') return false;
  const char* ds = std::getenv("B12X_DESCENT_LDS");
  return !(ds && ds[0] == '1');
}

// the end of a captured step whose keep ran the three-pass scan
__global__ void stepLogKernel(StepLogArgs g, uint32_t* scal) { logStep(g, scal, scal + 2); }

void enqueueStep(b12x_tree_s& T, b12x_scheme_s& S, uint8_t sd, uint32_t boundN, const uint32_t* countDev,
                 StepForm form, bool logStep = false) {
  SubdivideOp sop{};
  sop.in = T.view(T.cur);
  sop.out = T.view(1 - T.cur);
  sop.mask = sd;
  runScan(T, sop, boundN, T.totals, 0, countDev, 1);
  T.cur = 1 - T.cur;
  ImageArgs a{};
  a.dom = T.dom;
  a.map = S.map;
  a.grid = S.grid;
  a.key = T.key[T.cur];
  a.depth = T.depth[T.cur];
  a.srcKey = T.key[T.cur];
  a.srcDepth = T.depth[T.cur];
  a.nPoints = uint64_t(boundN) * 2 * S.grid.k;  // grid bound; the kernel reads the count
  a.nBoxesDev = T.scal + 0;
  a.mode = 0;
  a.hit = T.hit;
  if (form != STEP_EAGER) {
    launchImageFixed(T.dim, a, S.grid.kind, T.stream, kStepGraphBlocks);
  } else {
    a.samples = stageSamples(T, T.scal + 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (T.timing) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipEventRecord(e0, T.stream));
    }
    launchImage(T.dim, a, S.grid.kind, T.stream);
    if (T.timing) {
      HIPCHK(hipEventRecord(e1, T.stream));
      T.pending.push_back({e0, e1});
      T.imageLaunches += 1;
    }
  }
  KeepOp kop{};
  kop.in = T.view(T.cur);
  kop.out = T.view(1 - T.cur);
  kop.hit = T.hit;
  kop.mask = kop.hitFlag = kop.clearMask = uint8_t(F_HIT);
  kop.clearHit = 1;
  runScan(T, kop, boundN * 2 > 0 ? boundN * 2 : 1, T.totals, 2, T.scal + 0, 3);
  if (logStep) {  // the end of a captured step
    const StepLogArgs lg{T.scal + kScalStepLog, T.hscal + kScal, kop.out.key, kop.out.depth, kop.out.flags,
                         keyOf(0, 0)};
    hipLaunchKernelGGL(stepLogKernel, dim3(1), dim3(1), 0, T.stream, lg, T.scal);
    HIPCHK(hipGetLastError());
  }
  T.cur = 1 - T.cur;
}

// The one-block keep of a small step (keepSubdivideKernel) on the current half, in place; with
// hit == nullptr the batch's first subdivide, which also sets the batch's step count and log.
void launchKeepSubdivide(b12x_tree_s& T, uint8_t sd, uint8_t* hit, uint32_t stepsInit) {
  const KeepSubArgs k{T.view(T.cur), hit, T.scal, T.scal + kScalStepsLeft, stepsInit, T.scal + kScalStepLog,
                      T.hscal + kScal, sd, keyOf(0, 0)};
  hipLaunchKernelGGL(keepSubdivideKernel, dim3(1), dim3(kSmallBlock), 0, T.stream, k);
  HIPCHK(hipGetLastError());
}

// The captured small step: the image of the subdivided leaves (the current half, their count in
// scal[0]) and the keep that subdivides for the next step of the batch, both in place, so the
// step ends in the half it started in and replays with no host input.
void enqueueSmallStep(b12x_tree_s& T, b12x_scheme_s& S, uint8_t sd) {
  static_assert(kStepSmallLeaves <= kKeepSubMax, "a small step's leaves fit one block's keep");
  ImageArgs a{};
  a.dom = T.dom;
  a.map = S.map;
  a.grid = S.grid;
  a.key = a.srcKey = T.key[T.cur];
  a.depth = a.srcDepth = T.depth[T.cur];
  a.nPoints = uint64_t(kStepSmallLeaves) * S.grid.k;  // grid bound; the kernel reads the count
  a.nBoxesDev = T.scal + 0;
  a.mode = 0;
  a.hit = T.hit;
  launchImageFixed(T.dim, a, S.grid.kind, T.stream, kStepGraphBlocks);
  launchKeepSubdivide(T, sd, T.hit, 0);
}

// the leaf count where a captured step reads it (the keep of every step leaves it there too)
void stageLeafCount(b12x_tree_s& T) {
  T.hscal[kScalStepIn] = T.n;
  HIPCHK(hipMemcpyAsync(T.scal + 2, T.hscal + kScalStepIn, 4, hipMemcpyHostToDevice, T.stream));
}

// Image of every point of `src` located on the leaves (or on the root when depth0).
// mode 0 sets hit bytes, mode 1 writes targets to out.
// images of the source's test points [pointBase, pointBase + nPoints) (all of them by default);
// out receives them from index 0
void runImage(b12x_tree_s& T, b12x_scheme_s& S, const Source& src, bool depth0, int mode, uint32_t* out,
              const DepthMap* dm = nullptr, uint64_t* outKey = nullptr, uint64_t pointBase = 0,
              uint64_t nPoints = ~uint64_t(0)) {
  ImageArgs a{};
  a.outKey = outKey;
  if (dm) {  // search among the boxes at depth dm->d
    a.depthBox = dm->box;
    a.depthIn = dm->in;
    a.searchDepth = dm->d;
  }
  a.dom = T.dom;
  a.map = S.map;
  a.grid = S.grid;
  a.key = T.key[T.cur];
  a.depth = T.depth[T.cur];
  a.n = T.n;
  a.srcKey = src.key;
  a.srcDepth = src.depth;
  a.srcList = src.list;
  a.nPoints = nPoints == ~uint64_t(0) ? src.nBoxes * S.grid.k : nPoints;
  a.pointBase = pointBase;
  a.mode = mode;
  a.depth0 = depth0 ? 1 : 0;
  a.hit = T.hit;
  a.out = out;
  if (a.nPoints == 0) return;
  if (!depth0 && !dm && T.n > 0) a.samples = stageSamples(T, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (T.timing) {
    for (hipEvent_t* e : {&e0, &e1}) {
      if (!T.eventPool.empty()) {
        *e = T.eventPool.back();
        T.eventPool.pop_back();
      } else {
        HIPCHK(hipEventCreate(e));
      }
    }
    HIPCHK(hipEventRecord(e0, T.stream));
  }
  launchImage(T.dim, a, S.grid.kind, T.stream);
  if (T.timing) {
    HIPCHK(hipEventRecord(e1, T.stream));
    T.pending.push_back({e0, e1});
    T.imageLaunches += 1;
    T.imagePoints += a.nPoints;
    T.launchPoints.push_back(a.nPoints);
  }
}

// Fold completed image timings into the totals (called after the stream has synchronised).
// A device uint32 array that grows by doubling, keeping its contents: the outputs of passes
// whose sizes become known one pass at a time.
struct GrowBuf {
  uint32_t* p = nullptr;
  uint64_t cap = 0;
  GrowBuf() = default;
  GrowBuf(const GrowBuf&) = delete;
  GrowBuf& operator=(const GrowBuf&) = delete;
  ~GrowBuf() { dfree(p); }
  void reserve(uint64_t need, uint64_t used, hipStream_t st) {
    if (need <= cap) return;
    const uint64_t nc = std::max(need, 2 * cap);
    uint32_t* q = dalloc<uint32_t>(std::max<uint64_t>(nc, 1));
    if (used) HIPCHK(hipMemcpyAsync(q, p, used * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(p);
    p = q;
    cap = nc;
  }
  // to exactly `used` entries
  void shrink(uint64_t used, hipStream_t st) {
    if (cap <= used) return;
    uint32_t* q = dalloc<uint32_t>(std::max<uint64_t>(used, 1));
    if (used) HIPCHK(hipMemcpyAsync(q, p, used * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(p);
    p = q;
    cap = used;
  }
  uint32_t* release() {
    uint32_t* q = p;
    p = nullptr;
    cap = 0;
    return q;
  }
};

// Scratch bound of one transition-matrix pass: the tree's setting, else a third of the device's
// free memory (the reference sizes its first pass at a third of Memory<DEVICE>::getFreeBytes,
// ThrustSystem.hpp:64-68, ImplicitBoxTree.hpp:2941-2944)
uint64_t tmChunkBudget(const b12x_tree_s& T) {
  if (T.tmPassBytes) return T.tmPassBytes;
  size_t freeB = 0, totalB = 0;
  HIPCHK(hipMemGetInfo(&freeB, &totalB));
  return uint64_t(freeB) / 3;
}

thread_local uint32_t lastTmPasses = 0;  // passes of the calling thread's last transition matrix

void collectTimings(b12x_tree_s& T) {
  for (auto& p : T.pending) {
    HIPCHK(hipEventSynchronize(p.second));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.first, p.second));
    T.imageMs += ms;
    T.launchMs.push_back(ms);
    T.eventPool.push_back(p.first);
    T.eventPool.push_back(p.second);
  }
  T.pending.clear();
}

void checkScheme(b12x_tree_s& T, b12x_scheme_s* S) {
  if (!S) throw HipFail(B12X_E_ARG, "null scheme");
  if (S->dim != T.dim) throw HipFail(B12X_E_ARG, "scheme dimension differs from tree dimension");
}

// Flag operation over a selector (nullptr = all leaves); returns the reference's count.
// the flag operation on the leaves (selected by sel when set); its count lands in scal[0].
// withStats: the leaves' statistics before the operation land in scal[8..11] too.
void enqueueFlags(b12x_tree_s& T, uint8_t* sel, bool clearSel, int op, uint8_t a, uint8_t b, bool withStats) {
  zeroScal(T);
  if (withStats) enqueueLeafStats(T);
  hipLaunchKernelGGL(flagsKernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream, T.flags[T.cur], T.n, sel,
                     clearSel ? 1 : 0, op, a, b, T.scal);
  HIPCHK(hipGetLastError());
}

uint32_t applyFlags(b12x_tree_s& T, uint8_t* sel, bool clearSel, int op, uint8_t a, uint8_t b) {
  enqueueFlags(T, sel, clearSel, op, a, b, false);
  syncScal(T, 1);
  return T.hscal[0];
}

bool rootPred(int op, const LeafStats& s, uint8_t a) {  // countIfFlags at depth 0
  return op == 0 ? !areFlagsSet(s.flagsAnd, a) : (op == 1 ? isAnyFlagSet(s.flagsOr, a) : areFlagsSet(s.flagsAnd, a));
}

uint32_t stencilOp(b12x_tree_s& T, int op, const uint8_t* hostStencil, uint8_t a, uint8_t b, uint8_t d) {
  const DepthClass dc = classifyDepth(T, d);
  if (dc == DC_BEYOND) return 0;
  if (dc == DC_GENERAL) {  // the stencil has one entry per box at depth d
    DepthMap M;
    buildDepthMap(T, d, M);
    const uint32_t m = std::max<uint32_t>(M.count, 1);
    uint8_t* sel = dalloc<uint8_t>(m);
    if (hostStencil && M.count) HIPCHK(hipMemcpyAsync(sel, hostStencil, M.count, hipMemcpyHostToDevice, T.stream));
    else hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream, sel, uint64_t(m), uint8_t(1));
    const uint32_t r = depthFlagsOp(T, M, sel, op, a, b);
    dfree(sel);
    return r;
  }
  if (dc == DC_ROOT) {  // the count is the root's, from the statistics before the operation
    if (hostStencil && hostStencil[0] == 0) return 0;
    enqueueFlags(T, nullptr, false, op, a, b, /*withStats=*/true);
    syncScal(T, 12);
    return rootPred(op, leafStatsRead(T), a) ? 1u : 0u;
  }
  uint8_t* sel = nullptr;
  if (hostStencil) {
    sel = T.hit;  // hit bytes are all zero between operations; reuse them for the stencil
    HIPCHK(hipMemcpyAsync(sel, hostStencil, T.n, hipMemcpyHostToDevice, T.stream));
  }
  const uint32_t r = applyFlags(T, sel, sel != nullptr, op, a, b);
  if (sel) {
    hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream, sel, uint64_t(T.n),
                       uint8_t(0));
    HIPCHK(hipStreamSynchronize(T.stream));
  }
  return r;
}

}  // namespace

// ================================================================================================
// C ABI
// ================================================================================================
extern "C" {

const char* b12x_last_error(void) { return lastError().c_str(); }

int b12x_bench_copy(uint64_t bytes, int reps, double* ms) {
  return guarded([&]() -> int {
    if (!ms || reps < 1 || bytes < 16 || bytes % 16) return setErr(B12X_E_ARG, "bytes: a positive multiple of 16");
    if (int e = ensureDevice()) return e;
    const uint64_t n = bytes / 16;
    uint4* a = dalloc<uint4>(n);
    uint4* b = dalloc<uint4>(n);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const uint32_t grid = uint32_t(std::min<uint64_t>(ceilDiv(n, kBlock), 65536));
    hipLaunchKernelGGL(copyKernel, dim3(grid), dim3(kBlock), 0, nullptr, (const uint4*)a, b, n);  // first touch
    HIPCHK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(copyKernel, dim3(grid), dim3(kBlock), 0, nullptr, (const uint4*)a, b, n);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    return B12X_OK;
  });
}

int b12x_bench_fp64_clock(double* ghz, double* frac) {
  return guarded([&]() -> int {
    if (!ghz) return setErr(B12X_E_ARG, "null output");
    if (int e = ensureDevice()) return e;
    constexpr uint32_t kBlocks = 16384;
    constexpr int kIters = 2000;
    double* out = dalloc<double>(uint64_t(kBlocks) * kBlock);
    uint64_t* clk = dalloc<uint64_t>(2 * kBlocks);
    std::vector<uint64_t> h(2 * kBlocks);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(fp64ClockKernel, dim3(kBlocks), dim3(kBlock), 0, nullptr, out, clk, kIters, 0.999);
    HIPCHK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(fp64ClockKernel, dim3(kBlocks), dim3(kBlock), 0, nullptr, out, clk, kIters, 0.999);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    HIPCHK(hipMemcpy(h.data(), clk, sizeof(uint64_t) * 2 * kBlocks, hipMemcpyDeviceToHost));
    double cyc = 0.0, real = 0.0;
    for (uint32_t b = 0; b < kBlocks; ++b) {
      cyc += double(h[2 * b]);
      real += double(h[2 * b + 1]);
    }
    *ghz = real > 0.0 ? cyc / real * 0.1 : 0.0;  // the real-time counter runs at 100 MHz
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 16 FP64 wave instructions per iteration and wave, 4 cycles each on a 16-lane SIMD
    const double instr = double(kBlocks) * (kBlock / kWave) * kIters * 16.0;
    if (frac) *frac = (t > 0.0f && *ghz > 0.0) ? instr * 4.0 / (4.0 * cus * double(t) * 1e-3 * (*ghz) * 1e9) : 0.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(out);
    dfree(clk);
    return B12X_OK;
  });
}

int b12x_device_name(char* buf, int len) {
  return guarded([&]() -> int {
    if (int e = ensureDevice()) return e;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    std::snprintf(buf, size_t(len), "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return B12X_OK;
  });
}

int b12x_tree_create(int dim, const double* center, const double* radius, const uint8_t* sd, int nsd,
                     b12x_tree* out) {
  return guarded([&]() -> int {
    if (!out || !center || !radius) return setErr(B12X_E_ARG, "null argument");
    if (dim < 1 || dim > kMaxDim) return setErr(B12X_E_ARG, "dim must be 1..4");
    if (int e = ensureDevice()) return e;
    auto* T = new b12x_tree_s();
    T->dim = dim;
    T->sdEntries = sd ? kMaxDepth + 1 : kMaxDepth;
    if (!buildDomain(dim, center, radius, sd, nsd, T->dom)) {
      delete T;
      return setErr(B12X_E_ARG, "sdScheme has values >= dim");
    }
    try {
      streamCreate(&T->stream);
      T->scal = dalloc<uint32_t>(kScal);
      T->hscalBytes = kHostScal * sizeof(uint32_t);
      T->hscal = static_cast<uint32_t*>(ObjectPool::get().takePinned(T->hscalBytes));
      T->rootKey = dalloc<uint64_t>(1);
      T->rootDepth = dalloc<uint8_t>(1);
      hipLaunchKernelGGL(putRootKernel, dim3(1), dim3(1), 0, T->stream, T->rootKey, T->rootDepth, (uint8_t*)nullptr,
                         keyOf(0, 0));
      HIPCHK(hipGetLastError());
      insertRootIfEmpty(*T);
      HIPCHK(hipStreamSynchronize(T->stream));
    } catch (...) {
      delete T;
      throw;
    }
    *out = T;
    return B12X_OK;
  });
}

int b12x_tree_geometry(b12x_tree t, int* dim, double* center, double* radius, uint8_t* sd, int* nsd) {
  return guarded([&]() -> int {
    if (!t) return setErr(B12X_E_ARG, "null tree");
    if (dim) *dim = t->dim;
    for (int j = 0; j < t->dim; ++j) {
      if (center) center[j] = t->dom.center[j];
      if (radius) radius[j] = t->dom.radius[j];
    }
    if (sd)
      for (int i = 0; i < t->sdEntries; ++i) sd[i] = t->dom.sdScheme[i];
    if (nsd) *nsd = t->sdEntries;
    return B12X_OK;
  });
}

int b12x_tree_destroy(b12x_tree t) {
  delete t;
  return B12X_OK;
}

// setCenter / setRadius / setSdScheme (ImplicitBoxTree.hpp:3464-3552): the leaves stay, the
// domain they are read in changes; captured subdivision steps hold the old domain and are rebuilt
int b12x_tree_set_domain(b12x_tree t, const double* center, const double* radius, const uint8_t* sd, int nsd,
                         int setSd) {
  return guarded([&]() -> int {
    if (!t) return setErr(B12X_E_ARG, "null tree");
    b12x_tree_s& T = *t;
    std::vector<double> c(T.dom.center, T.dom.center + T.dim), r(T.dom.radius, T.dom.radius + T.dim);
    if (center) c.assign(center, center + T.dim);
    if (radius) {
      for (int j = 0; j < T.dim; ++j)
        if (!(radius[j] >= 0.0)) return setErr(B12X_E_ARG, "Input has negative values. Radius vector remains unchanged.");
      r.assign(radius, radius + T.dim);
    }
    std::vector<uint8_t> keep;
    const uint8_t* sdUse = nullptr;
    int nUse = 0, entries = T.sdEntries;
    if (setSd) {
      for (int i = 0; i < nsd; ++i)
        if (sd[i] >= T.dim) return setErr(B12X_E_ARG, "Input has values >= DIM. SdScheme remains unchanged.");
      if (nsd > 0) {
        sdUse = sd;
        nUse = nsd;
      }
      entries = nsd > 0 ? kMaxDepth + 1 : kMaxDepth;  // a given scheme: 65 entries, the default 64
    } else if (T.sdEntries != kMaxDepth) {
      keep.assign(T.dom.sdScheme, T.dom.sdScheme + T.sdEntries);
      sdUse = keep.data();
      nUse = int(keep.size());
    }
    Domain D{};
    if (!buildDomain(T.dim, c.data(), r.data(), sdUse, nUse, D)) return setErr(B12X_E_ARG, "sdScheme has values >= dim");
    T.dom = D;
    T.sdEntries = entries;
    ++T.bufGen;  // the captured steps carry the domain as a kernel argument
    return B12X_OK;
  });
}

// areAllLeavesOnDepth (ImplicitBoxTree.hpp:589-609)
int b12x_all_leaves_on_depth(b12x_tree t, uint8_t d, int* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    if (d == kLeafDepth) *out = 1;
    else if (d == kUnsubDepth) *out = 0;
    else *out = classifyDepth(*t, d) == DC_LEAF || (d == 0 && t->n == 1 && leafStats(*t).maxDepth == 0) ? 1 : 0;
    return B12X_OK;
  });
}

// hasSameBoxes (ImplicitBoxTree.hpp:1299-1381): dimension, centre, radius, the leaves and the
// subdivision scheme up to depth(); the reference compares this tree's leaves with the first
// leaves of the other, so a longer other tree with the same first leaves counts as the same
int b12x_has_same_boxes(b12x_tree a, b12x_tree b, int* out) {
  return guarded([&]() -> int {
    if (!a || !b || !out) return setErr(B12X_E_ARG, "null argument");
    *out = 0;
    if (a->dim != b->dim) return B12X_OK;
    for (int j = 0; j < a->dim; ++j)
      if (!(a->dom.center[j] == b->dom.center[j]) || !(a->dom.radius[j] == b->dom.radius[j])) return B12X_OK;
    if (b->n < a->n) return B12X_OK;  // the reference would read past the other tree's leaves
    const uint8_t dA = leafStats(*a).maxDepth;
    for (int i = 0; i < int(dA) && i <= kMaxDepth; ++i)
      if (a->dom.sdScheme[i] != b->dom.sdScheme[i]) return B12X_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemsetAsync(a->scal + 24, 0, 4, a->stream));
    hipLaunchKernelGGL(sameLeavesKernel, dim3(ceilDiv(a->n, kBlock)), dim3(kBlock), 0, a->stream, a->view(a->cur),
                       b->view(b->cur), a->n, a->scal + 24);
    HIPCHK(hipGetLastError());
    syncScal(*a, 25);
    *out = a->hscal[24] == 0 ? 1 : 0;
    return B12X_OK;
  });
}

// getNHostBufferBytes / setNHostBufferBytes (ImplicitBoxTree.hpp:3572-3583): kept and written to
// files; every tree here lives in HBM whole, so the value does not limit anything
int b12x_get_host_buffer_bytes(b12x_tree t, uint64_t* out) {
  if (!t || !out) return setErr(B12X_E_ARG, "null argument");
  *out = t->hostBufferBytes;
  return B12X_OK;
}

int b12x_set_host_buffer_bytes(b12x_tree t, uint64_t v) {
  if (!t) return setErr(B12X_E_ARG, "null tree");
  t->hostBufferBytes = v;
  return B12X_OK;
}

int b12x_set_tm_pass_bytes(b12x_tree t, uint64_t bytes) {
  if (!t) return setErr(B12X_E_ARG, "null tree");
  t->tmPassBytes = bytes;
  return B12X_OK;
}

int b12x_get_tm_pass_bytes(b12x_tree t, uint64_t* bytes) {
  if (!t || !bytes) return setErr(B12X_E_ARG, "null argument");
  *bytes = t->tmPassBytes;
  return B12X_OK;
}

int b12x_tm_last_passes(uint32_t* passes) {
  if (!passes) return setErr(B12X_E_ARG, "null argument");
  *passes = lastTmPasses;
  return B12X_OK;
}

int b12x_size(b12x_tree t, uint32_t* out) {
  if (!t || !out) return setErr(B12X_E_ARG, "null argument");
  *out = t->n;
  return B12X_OK;
}

int b12x_count(b12x_tree t, uint8_t d, uint32_t* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    // ImplicitBoxTree.hpp:1060-1095 for the root, the leaves and a depth shared by all leaves
    const DepthClass dc = classifyDepth(*t, d);
    if (dc == DC_ROOT) *out = 1;
    else if (dc == DC_LEAF) *out = t->n;
    else if (dc == DC_BEYOND) *out = 0;
    else {
      DepthMap M;
      buildDepthMap(*t, d, M);
      *out = M.count;
    }
    return B12X_OK;
  });
}

int b12x_depth(b12x_tree t, uint8_t* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    *out = leafStats(*t).maxDepth;
    return B12X_OK;
  });
}

int b12x_tree_download(b12x_tree t, uint64_t* bits, uint8_t* depth, uint8_t* flags) {
  return guarded([&]() -> int {
    if (!t) return setErr(B12X_E_ARG, "null tree");
    std::vector<uint64_t> k(t->n);
    std::vector<uint8_t> d(t->n);
    HIPCHK(hipMemcpyAsync(k.data(), t->key[t->cur], size_t(t->n) * 8, hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipMemcpyAsync(d.data(), t->depth[t->cur], t->n, hipMemcpyDeviceToHost, t->stream));
    if (flags) HIPCHK(hipMemcpyAsync(flags, t->flags[t->cur], t->n, hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    for (uint32_t i = 0; i < t->n; ++i) {
      if (bits) bits[i] = bitsOf(k[i], d[i]);
      if (depth) depth[i] = d[i];
    }
    return B12X_OK;
  });
}

int b12x_tree_upload(b12x_tree t, uint32_t n, const uint64_t* bits, const uint8_t* depth, const uint8_t* flags) {
  return guarded([&]() -> int {
    if (!t || (n && (!bits || !depth || !flags))) return setErr(B12X_E_ARG, "null argument");
    struct E {
      uint64_t k;
      uint8_t d, f;
    };
    std::vector<E> v(n);
    for (uint32_t i = 0; i < n; ++i) {
      if (depth[i] > kMaxDepth) return setErr(B12X_E_ARG, "leaf depth exceeds 64");
      v[i] = {keyOf(bits[i], depth[i]), depth[i], flags[i]};
    }
    // isStrictlyPreceding order (Box.hpp:122-135): ascending key, deeper first on ties
    std::stable_sort(v.begin(), v.end(), [](const E& a, const E& b) { return a.k < b.k || (a.k == b.k && b.d < a.d); });
    std::vector<uint64_t> k(n);
    std::vector<uint8_t> d(n), f(n);
    for (uint32_t i = 0; i < n; ++i) {
      k[i] = v[i].k;
      d[i] = v[i].d;
      f[i] = v[i].f;
    }
    t->n = 0;
    ensureCap(*t, std::max<uint32_t>(n, 1));
    HIPCHK(hipMemcpyAsync(t->key[t->cur], k.data(), size_t(n) * 8, hipMemcpyHostToDevice, t->stream));
    HIPCHK(hipMemcpyAsync(t->depth[t->cur], d.data(), n, hipMemcpyHostToDevice, t->stream));
    HIPCHK(hipMemcpyAsync(t->flags[t->cur], f.data(), n, hipMemcpyHostToDevice, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    t->n = n;
    ++t->leafGen;
    insertRootIfEmpty(*t);
    return B12X_OK;
  });
}

int b12x_subdivide(b12x_tree t, uint8_t f, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    ensureCap(T, uint64_t(T.n) * 2);
    zeroScal(T);
    SubdivideOp op{};
    op.in = T.view(T.cur);
    op.out = T.view(1 - T.cur);
    op.mask = f;
    runScan(T, op, T.n, T.totals, 0, nullptr, 1);
    syncScal(T, 2);
    T.n = T.hscal[0];
    T.cur = 1 - T.cur;
    ++T.leafGen;
    *ret = T.hscal[1];
    return B12X_OK;
  });
}

int b12x_unsubdivide(b12x_tree t, uint8_t f, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    zeroScal(T);
    const int o = 1 - T.cur;
    UnsubMarkKernelArgs a{T.view(T.cur), T.n, f, T.key[o], T.depth[o], T.scal + 1};
    hipLaunchKernelGGL(unsubMarkKernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream, a);
    // lifted keys/depths sit in the other buffer; merge back into the current one
    UnsubMergeOp op{};
    op.key = T.key[o];
    op.depth = T.depth[o];
    op.flags = T.flags[T.cur];
    op.n = T.n;
    // the merge writes keys/depths in place over the lifted arrays: positions never exceed i,
    // but other blocks may still read them, so go through the hit-sized staging instead
    uint64_t* k2 = dalloc<uint64_t>(T.n);
    uint8_t* d2 = dalloc<uint8_t>(T.n);
    uint8_t* f2 = dalloc<uint8_t>(T.n);
    op.out = {k2, d2, f2};
    runScan(T, op, T.n, T.totals, 0);
    syncScal(T, 2);
    const uint32_t m = T.hscal[0];
    HIPCHK(hipMemcpyAsync(T.key[T.cur], k2, size_t(m) * 8, hipMemcpyDeviceToDevice, T.stream));
    HIPCHK(hipMemcpyAsync(T.depth[T.cur], d2, m, hipMemcpyDeviceToDevice, T.stream));
    HIPCHK(hipMemcpyAsync(T.flags[T.cur], f2, m, hipMemcpyDeviceToDevice, T.stream));
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(k2);
    dfree(d2);
    dfree(f2);
    *ret = T.hscal[1] - (T.n - m);
    T.n = m;
    ++T.leafGen;
    return B12X_OK;
  });
}

int b12x_remove(b12x_tree t, uint8_t f, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    zeroScal(T);
    KeepOp op{};
    op.in = T.view(T.cur);
    op.out = T.view(1 - T.cur);
    op.hit = nullptr;
    op.mask = f;
    runScan(T, op, T.n, T.totals, 0);
    syncScal(T, 1);
    const uint32_t m = T.hscal[0];
    *ret = T.n - m;
    T.n = m;
    T.cur = 1 - T.cur;
    ++T.leafGen;
    insertRootIfEmpty(T);
    return B12X_OK;
  });
}

}  // extern "C"

namespace {

__global__ void iotaKernel(uint32_t* p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = uint32_t(i);
}

// Insertion (ImplicitBoxTree.hpp:1613-1878) of the P point keys pk (maximal depth, or invalid)
// whose search results at depth d are sv, after the depth map M of d (general depths) was built:
// marks the points to insert and the hit boxes, sets flagsForHit, then merges the new boxes into
// the leaves (sort, runs on one path reduced, neighbours on one path collapsed). resHost (may be
// null) receives isToInsert per point.
void insertCore(b12x_tree_s& T, uint64_t P, const uint64_t* pk, const uint32_t* sv, uint8_t d, DepthClass dc,
                const DepthMap* M, uint8_t fIns, uint8_t fHit, uint8_t* resHost) {
  const int form = d == kLeafDepth ? 0 : (dc == DC_BEYOND ? 2 : 1);
  const uint32_t n = T.n;
  DevScope ds;  // every buffer below except the radix sorts' own (they swap and free theirs)
  uint8_t* ins = ds.alloc<uint8_t>(P);
  uint32_t* insU32 = ds.alloc<uint32_t>(P + 1);
  hipLaunchKernelGGL(insertMarkKernel, dim3(ceilDiv(P, kBlock)), dim3(kBlock), 0, T.stream, pk, sv, P, form,
                     form == 0 ? 0 : int(kMaxDepth) - int(d), (const uint8_t*)T.depth[T.cur], ins, insU32, T.hit,
                     (uint32_t*)nullptr);
  HIPCHK(hipGetLastError());
  // flagsForHit on the hit boxes; the hit bytes end zero
  if (form != 2) {
    if (form == 0 || dc == DC_LEAF) {
      zeroScal(T);
      hipLaunchKernelGGL(flagsKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.flags[T.cur], n, T.hit, 1,
                         0, fHit, uint8_t(0), T.scal + 0);
    } else {
      (void)depthFlagsOp(T, *M, T.hit, 0, fHit, 0);
      hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(std::max<uint32_t>(M->count, 1), kBlock)), dim3(kBlock), 0,
                         T.stream, T.hit, uint64_t(M->count), uint8_t(0));
    }
    HIPCHK(hipGetLastError());
  }
  // the points may outnumber the leaves: the scan over them has scratch of its own
  uint32_t* ptTotals = ds.alloc<uint32_t>(uint64_t(ceilDiv(P, kTile)) + 16);
  const uint32_t m = scanArray(T.stream, insU32, P, ptTotals, T.scal + 4, T.hscal + 4);
  if (resHost) HIPCHK(hipMemcpyAsync(resHost, ins, P, hipMemcpyDeviceToHost, T.stream));
  if (m == 0) {
    HIPCHK(hipStreamSynchronize(T.stream));
    return;
  }
  // the new boxes, sorted (comparable key ascending, deeper first on equal keys, stable), then
  // merged into the leaves, which are sorted already
  const uint64_t N = uint64_t(n) + m;
  ensureCap(T, N);
  uint64_t* newK = ds.alloc<uint64_t>(m);
  uint8_t* newD = ds.alloc<uint8_t>(m);
  uint8_t* newF = ds.alloc<uint8_t>(m);
  hipLaunchKernelGGL(insertNewKernel, dim3(ceilDiv(P, kBlock)), dim3(kBlock), 0, T.stream, pk,
                     (const uint32_t*)insU32, m, P, form == 0 ? 1 : 0, d,
                     (const uint64_t*)T.key[T.cur], (const uint8_t*)T.depth[T.cur], n, fIns, newK, newD, newF);
  HIPCHK(hipGetLastError());
  uint64_t* nsK = ds.alloc<uint64_t>(m);
  uint8_t* nsD = ds.alloc<uint8_t>(m);
  uint8_t* nsF = ds.alloc<uint8_t>(m);
  if (m <= kSmallSortMax) {
    hipLaunchKernelGGL(smallBoxSortKernel, dim3(1), dim3(kSmallSortThreads), 0, T.stream, (const uint64_t*)newK,
                       (const uint8_t*)newD, (const uint8_t*)newF, m, nsK, nsD, nsF);
  } else {
    // two stable LSD radix sorts by the key's low and high words (equal keys are equal boxes)
    uint32_t* idx = dalloc<uint32_t>(m);
    uint32_t* k32 = dalloc<uint32_t>(m);
    uint32_t* dummy = dalloc<uint32_t>(m);
    hipLaunchKernelGGL(iotaKernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream, idx, uint64_t(m));
    for (int part = 1; part < 3; ++part) {
      hipLaunchKernelGGL(boxSortKeyKernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream,
                         (const uint64_t*)newK, (const uint8_t*)newD, (const uint32_t*)idx, uint64_t(m), part, k32);
      HIPCHK(hipGetLastError());
      radixSort3(T.stream, k32, idx, dummy, m, 32, T.hscal + 5);
    }
    hipLaunchKernelGGL(gatherBoxesKernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream,
                       (const uint64_t*)newK, (const uint8_t*)newD, (const uint8_t*)newF, (const uint32_t*)idx,
                       uint64_t(m), nsK, nsD, nsF);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(idx);
    dfree(k32);
    dfree(dummy);
  }
  HIPCHK(hipGetLastError());
  uint64_t* sK = ds.alloc<uint64_t>(N);
  uint8_t* sD = ds.alloc<uint8_t>(N);
  uint8_t* sF = ds.alloc<uint8_t>(N);
  hipLaunchKernelGGL(mergeNewKernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream, (const uint64_t*)nsK,
                     (const uint8_t*)nsD, (const uint8_t*)nsF, m, (const uint64_t*)T.key[T.cur],
                     (const uint8_t*)T.depth[T.cur], n, sK, sD, sF);
  hipLaunchKernelGGL(mergeLeafKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream,
                     (const uint64_t*)T.key[T.cur], (const uint8_t*)T.depth[T.cur], (const uint8_t*)T.flags[T.cur], n,
                     (const uint64_t*)nsK, (const uint8_t*)nsD, m, sK, sD, sF);
  HIPCHK(hipGetLastError());
  // runs on one path, reduced
  uint32_t* run = ds.alloc<uint32_t>(N);
  uint8_t* runStart = ds.alloc<uint8_t>(N);
  hipLaunchKernelGGL(pathRunKernel, dim3(ceilDiv(N, kBlock)), dim3(kBlock), 0, T.stream, (const uint64_t*)sK,
                     (const uint8_t*)sD, uint32_t(N), run, runStart);
  HIPCHK(hipGetLastError());
  const uint32_t R = scanArray(T.stream, run, N, T.totals, T.scal + 6, T.hscal + 6);
  hipLaunchKernelGGL(depthBoxKernel, dim3(ceilDiv(N, kBlock)), dim3(kBlock), 0, T.stream, run,
                     (const uint8_t*)runStart, uint32_t(N));
  unsigned long long* rb = ds.alloc<unsigned long long>(R);
  uint32_t* rd = ds.alloc<uint32_t>(R);
  uint32_t* rf = ds.alloc<uint32_t>(R);
  HIPCHK(hipMemsetAsync(rb, 0, size_t(R) * 8, T.stream));
  HIPCHK(hipMemsetAsync(rd, 0, size_t(R) * 4, T.stream));
  HIPCHK(hipMemsetAsync(rf, 0, size_t(R) * 4, T.stream));
  hipLaunchKernelGGL(pathReduceKernel, dim3(ceilDiv(N, kBlock)), dim3(kBlock), 0, T.stream, (const uint64_t*)sK,
                     (const uint8_t*)sD, (const uint8_t*)sF, (const uint32_t*)run, uint32_t(N), rb, rd, rf);
  uint32_t* keep = ds.alloc<uint32_t>(R);
  hipLaunchKernelGGL(pathKeepKernel, dim3(ceilDiv(R, kBlock)), dim3(kBlock), 0, T.stream,
                     (const unsigned long long*)rb, (const uint32_t*)rd, R, keep);
  HIPCHK(hipGetLastError());
  const uint32_t K = scanArray(T.stream, keep, R, T.totals, T.scal + 7, T.hscal + 7);
  hipLaunchKernelGGL(pathWriteKernel, dim3(ceilDiv(R, kBlock)), dim3(kBlock), 0, T.stream,
                     (const unsigned long long*)rb, (const uint32_t*)rd, (const uint32_t*)rf, (const uint32_t*)keep,
                     R, T.view(1 - T.cur));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(T.stream));
  T.cur = 1 - T.cur;
  T.n = K;
  ++T.leafGen;
}

}  // namespace

extern "C" {

// ImplicitBoxTree.hpp:1215-1262 deleteDepth
int b12x_delete_depth(b12x_tree t, uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    if (d == 0) {
      T.n = 0;
      insertRootIfEmpty(T);
      *ret = 1;
      return B12X_OK;
    }
    const DepthClass dc = classifyDepth(T, d);
    if (d == kLeafDepth || d == kUnsubDepth || dc == DC_BEYOND) {
      *ret = 0;
      return B12X_OK;
    }
    uint32_t nb = T.n;  // count(d)
    if (dc == DC_GENERAL) {
      DepthMap M;
      buildDepthMap(T, d, M);
      nb = M.count;
    }
    const uint32_t n = T.n;
    const uint8_t d1 = uint8_t(d - 1);
    DevScope ds;
    uint32_t* run = ds.alloc<uint32_t>(n);
    uint8_t* runStart = ds.alloc<uint8_t>(n);
    hipLaunchKernelGGL(truncRunKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.view(T.cur), n, d1, run,
                       runStart);
    HIPCHK(hipGetLastError());
    const uint32_t m = scanArray(T.stream, run, n, T.totals, T.scal + 4, T.hscal + 4);
    hipLaunchKernelGGL(depthBoxKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, run,
                       (const uint8_t*)runStart, n);
    uint32_t* acc = ds.alloc<uint32_t>(std::max<uint32_t>(m, 1));
    hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(std::max<uint32_t>(m, 1), kBlock)), dim3(kBlock), 0, T.stream, acc,
                       uint64_t(std::max<uint32_t>(m, 1)), 0xFFu);
    hipLaunchKernelGGL(depthCombineKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream,
                       (const uint8_t*)T.flags[T.cur], n, (const uint32_t*)run, (const uint8_t*)nullptr, 0, acc);
    hipLaunchKernelGGL(truncWriteKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.view(T.cur),
                       T.view(1 - T.cur), n, d1, (const uint8_t*)runStart, (const uint32_t*)run,
                       (const uint32_t*)acc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(T.stream));
    T.cur = 1 - T.cur;
    T.n = m;
    ++T.leafGen;
    *ret = nb;
    return B12X_OK;
  });
}

// ImplicitBoxTree.hpp:1613-1892 insertByIterators(points of (dIt, check, grid, map), depth,
// flagsForInserted, flagsForHit)
int b12x_insert_by_image(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, uint8_t d, uint8_t fIns,
                         uint8_t fHit, uint8_t* res) {
  return guarded([&]() -> int {
    if (!t) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const Source src = makeSource(T, dIt, check);
    const uint64_t P = src.nBoxes * s->grid.k;
    if (P == 0) return B12X_OK;
    if (P > 0xFFFFFFFFull) return setErr(B12X_E_UNSUPPORTED, "more than 2^32 points in one insertion");
    if (d != 0 && d != kLeafDepth && d > kMaxDepth) {  // :1872-1876
      if (res) std::memset(res, 0, P);
      return B12X_OK;
    }
    const DepthClass dc = d == 0 ? DC_ROOT : classifyDepth(T, d);
    DepthMap M;
    if (dc == DC_GENERAL) buildDepthMap(T, d, M);
    uint64_t* pk = dalloc<uint64_t>(P);
    uint32_t* sv = dalloc<uint32_t>(P);
    // the images: point keys and their search at d (leaf search when d is deeper than the tree)
    runImage(T, *s, src, false, 1, sv, dc == DC_GENERAL ? &M : nullptr, pk);
    if (d == 0) {  // :1621-1636 nothing inserted; flagsForHit on every leaf if a point is valid
      std::vector<uint64_t> h(P);
      HIPCHK(hipMemcpyAsync(h.data(), pk, P * 8, hipMemcpyDeviceToHost, T.stream));
      HIPCHK(hipStreamSynchronize(T.stream));
      bool any = false;
      for (uint64_t v : h) any = any || v != kInvalidKey;
      if (any && fHit) stencilOp(T, 0, nullptr, fHit, 0, kLeafDepth);
      if (res) std::memset(res, 0, P);
    } else {
      insertCore(T, P, pk, sv, d, dc, dc == DC_GENERAL ? &M : nullptr, fIns, fHit, res);
    }
    dfree(pk);
    dfree(sv);
    return B12X_OK;
  });
}

// ImplicitBoxTree.hpp:1894-1930 insertByPoints (the point boxes carry flagsForInserted)
int b12x_insert_by_points(b12x_tree t, const double* pts, uint64_t P, uint8_t d, uint8_t fIns, uint8_t fHit,
                          uint8_t* res) {
  return guarded([&]() -> int {
    if (!t || (P && !pts)) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    if (P == 0) return B12X_OK;
    if (P > 0xFFFFFFFFull) return setErr(B12X_E_UNSUPPORTED, "more than 2^32 points in one insertion");
    if (d != 0 && d != kLeafDepth && d > kMaxDepth) {
      if (res) std::memset(res, 0, P);
      return B12X_OK;
    }
    const DepthClass dc = d == 0 ? DC_ROOT : classifyDepth(T, d);
    DepthMap M;
    if (dc == DC_GENERAL) buildDepthMap(T, d, M);
    double* dp = dalloc<double>(P * T.dim);
    uint64_t* pk = dalloc<uint64_t>(P);
    uint32_t* sv = dalloc<uint32_t>(P);
    HIPCHK(hipMemcpyAsync(dp, pts, P * T.dim * sizeof(double), hipMemcpyHostToDevice, T.stream));
    const uint32_t nb = ceilDiv(P, kBlock);
    const uint32_t* mb = dc == DC_GENERAL ? M.box : nullptr;
    const uint8_t* mi = dc == DC_GENERAL ? M.in : nullptr;
    const uint8_t sd = dc == DC_GENERAL ? d : 0;
    switch (T.dim) {
      case 1: hipLaunchKernelGGL(pointsSearchKernel<1>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, P, T.key[T.cur], T.depth[T.cur], T.n, 0, sv, mb, mi, sd, pk); break;
      case 2: hipLaunchKernelGGL(pointsSearchKernel<2>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, P, T.key[T.cur], T.depth[T.cur], T.n, 0, sv, mb, mi, sd, pk); break;
      case 3: hipLaunchKernelGGL(pointsSearchKernel<3>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, P, T.key[T.cur], T.depth[T.cur], T.n, 0, sv, mb, mi, sd, pk); break;
      default: hipLaunchKernelGGL(pointsSearchKernel<4>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, P, T.key[T.cur], T.depth[T.cur], T.n, 0, sv, mb, mi, sd, pk); break;
    }
    HIPCHK(hipGetLastError());
    if (d == 0) {
      std::vector<uint64_t> h(P);
      HIPCHK(hipMemcpyAsync(h.data(), pk, P * 8, hipMemcpyDeviceToHost, T.stream));
      HIPCHK(hipStreamSynchronize(T.stream));
      bool any = false;
      for (uint64_t v : h) any = any || v != kInvalidKey;
      if (any && fHit) stencilOp(T, 0, nullptr, fHit, 0, kLeafDepth);
      if (res) std::memset(res, 0, P);
    } else {
      insertCore(T, P, pk, sv, d, dc, dc == DC_GENERAL ? &M : nullptr, fIns, fHit, res);
    }
    dfree(dp);
    dfree(pk);
    dfree(sv);
    return B12X_OK;
  });
}

// continuation.m:13-44 for one depth: deleteDepth(0); insert the points' boxes at depth with INS;
// then while count(depth) grows: INS -> EXPD at depth, insert the boxes hit by the images of the
// EXPD boxes with INS, unset EXPD on the leaves. The per-point results of the insertions are not
// kept, and where every leaf lies on `depth` the flag operations do not read their counts back.
int b12x_continuation(b12x_tree t, b12x_scheme s, uint8_t d, const double* pts, uint64_t P, uint32_t* counts,
                      int maxCounts, int* nCounts) {
  return guarded([&]() -> int {
    if (!t || !s || (P && !pts) || !nCounts || (maxCounts > 0 && !counts)) return setErr(B12X_E_ARG, "bad argument");
    b12x_tree_s& T = *t;
    constexpr uint8_t kIns = 0x02, kExpd = 0x04;
    auto flagsAt = [&](int op, uint8_t a, uint8_t b, uint8_t depth) {
      if (classifyDepth(T, depth) == DC_LEAF) {  // every leaf on `depth`: no count read back
        hipLaunchKernelGGL(flagsKernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream, T.flags[T.cur], T.n,
                           (uint8_t*)nullptr, 0, op, a, b, T.scal + 26);  // a count nobody reads
        HIPCHK(hipGetLastError());
      } else {
        (void)stencilOp(T, op, nullptr, a, b, depth);
      }
    };
    auto countAt = [&]() -> uint32_t {
      uint32_t c = 0;
      if (int e = b12x_count(t, d, &c)) throw HipFail(e, lastError());
      return c;
    };
    int k = 0;
    auto record = [&](uint32_t c) {
      if (k < maxCounts) counts[k] = c;
      ++k;
    };
    uint32_t r = 0;
    if (int e = b12x_delete_depth(t, 0, &r)) return e;
    if (int e = b12x_insert_by_points(t, pts, P, d, kIns, 0, nullptr)) return e;
    uint32_t nb0 = 0, nb1 = countAt();
    record(nb1);
    while (nb1 > nb0) {
      flagsAt(2, kIns, kExpd, d);
      if (int e = b12x_insert_by_image(t, s, d, kExpd, d, kIns, 0, nullptr)) return e;
      flagsAt(1, kExpd, 0, kLeafDepth);
      nb0 = nb1;
      nb1 = countAt();
      record(nb1);
    }
    HIPCHK(hipStreamSynchronize(T.stream));
    *nCounts = k;
    return B12X_OK;
  });
}

// ImplicitBoxTree.hpp:611-784 boxes(depth) and :1273-1283 enumerate()
int b12x_boxes(b12x_tree t, uint8_t d, double* out, uint64_t* nOut) {
  return guarded([&]() -> int {
    if (!t || !nOut) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    const int D = T.dim;
    const uint64_t W = uint64_t(2 * D + 1);
    const DepthClass dc = classifyDepth(T, d);
    if (dc == DC_BEYOND) {
      *nOut = 0;
      return B12X_OK;
    }
    if (dc == DC_ROOT) {  // the surrounding box, flags ANDed over all leaves
      *nOut = 1;
      if (out) {
        const LeafStats st = leafStats(T);
        for (int j = 0; j < D; ++j) {
          out[j] = T.dom.center[j];
          out[D + j] = T.dom.radius[j];
        }
        out[2 * D] = double(st.flagsAnd);
      }
      return B12X_OK;
    }
    DepthMap M;
    uint32_t* acc = nullptr;
    uint64_t nb = T.n;
    if (dc == DC_GENERAL) {
      buildDepthMap(T, d, M);
      nb = M.count;
    }
    *nOut = nb;
    if (!out || nb == 0) return B12X_OK;
    if (dc == DC_GENERAL) {
      acc = dalloc<uint32_t>(nb);
      hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(nb, kBlock)), dim3(kBlock), 0, T.stream, acc, nb, 0xFFu);
      hipLaunchKernelGGL(depthCombineKernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream,
                         (const uint8_t*)T.flags[T.cur], T.n, (const uint32_t*)M.box, (const uint8_t*)M.in, 0, acc);
    }
    double* dout = dalloc<double>(nb * W);
    const uint8_t* st = dc == DC_GENERAL ? M.start : nullptr;
    const dim3 g(ceilDiv(T.n, kBlock)), bl(kBlock);
    switch (D) {
      case 1: hipLaunchKernelGGL(boxesKernel<1>, g, bl, 0, T.stream, T.dom, T.view(T.cur), T.n, d, st, (const uint32_t*)M.box, (const uint32_t*)acc, dout); break;
      case 2: hipLaunchKernelGGL(boxesKernel<2>, g, bl, 0, T.stream, T.dom, T.view(T.cur), T.n, d, st, (const uint32_t*)M.box, (const uint32_t*)acc, dout); break;
      case 3: hipLaunchKernelGGL(boxesKernel<3>, g, bl, 0, T.stream, T.dom, T.view(T.cur), T.n, d, st, (const uint32_t*)M.box, (const uint32_t*)acc, dout); break;
      default: hipLaunchKernelGGL(boxesKernel<4>, g, bl, 0, T.stream, T.dom, T.view(T.cur), T.n, d, st, (const uint32_t*)M.box, (const uint32_t*)acc, dout); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, nb * W * sizeof(double), hipMemcpyDeviceToHost, T.stream));
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(dout);
    dfree(acc);
    return B12X_OK;
  });
}

int b12x_enumerate(b12x_tree t, uint32_t* out, int* nOut) {
  return guarded([&]() -> int {
    if (!t || !nOut) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    const int dmax = leafStats(T).maxDepth;
    *nOut = dmax + 1;
    if (!out) return B12X_OK;
    for (int d = 0; d <= dmax; ++d) {
      const DepthClass dc = classifyDepth(T, uint8_t(d));
      if (dc == DC_ROOT) out[d] = 1;
      else if (dc == DC_LEAF) out[d] = T.n;
      else if (dc == DC_BEYOND) out[d] = 0;
      else {
        DepthMap M;
        buildDepthMap(T, uint8_t(d), M);
        out[d] = M.count;
      }
    }
    return B12X_OK;
  });
}

int b12x_set_flags(b12x_tree t, uint8_t f, uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    *ret = stencilOp(*t, 0, nullptr, f, 0, d);
    return B12X_OK;
  });
}

int b12x_unset_flags(b12x_tree t, uint8_t f, uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    *ret = stencilOp(*t, 1, nullptr, f, 0, d);
    return B12X_OK;
  });
}

int b12x_change_flags(b12x_tree t, uint8_t fu, uint8_t fs, uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret) return setErr(B12X_E_ARG, "null argument");
    *ret = stencilOp(*t, 2, nullptr, fu, fs, d);
    return B12X_OK;
  });
}

int b12x_flags_by_stencil(b12x_tree t, int op, const uint8_t* st, uint8_t f, uint8_t f2, uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret || !st || op < 0 || op > 2) return setErr(B12X_E_ARG, "bad argument");
    *ret = stencilOp(*t, op, st, f, f2, d);
    return B12X_OK;
  });
}

int b12x_flags_by_energy_band(b12x_tree t, double emin, double emax, int op, uint8_t f, uint8_t f2, uint8_t d,
                              uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret || op < 0 || op > 2) return setErr(B12X_E_ARG, "bad argument");
    b12x_tree_s& T = *t;
    if (T.dim != 4) return setErr(B12X_E_ARG, "energy band: 4-dimensional trees only");
    if (classifyDepth(T, d) != DC_LEAF) throw HipFail(B12X_E_UNSUPPORTED, "energy band: leaf depth only");
    hipLaunchKernelGGL(energyBandKernel, dim3(ceilDiv(T.n, kBlock)), dim3(kBlock), 0, T.stream, T.dom,
                       (const uint64_t*)T.key[T.cur], (const uint8_t*)T.depth[T.cur], T.n, emin, emax, T.hit);
    HIPCHK(hipGetLastError());
    *ret = applyFlags(T, T.hit, true, op, f, f2);  // consumes and clears the selection
    return B12X_OK;
  });
}

int b12x_num_points(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, uint64_t* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    checkScheme(*t, s);
    *out = makeSource(*t, dIt, check).nBoxes * s->grid.k;
    return B12X_OK;
  });
}

int b12x_flags_by_image(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, int op, uint8_t f, uint8_t f2,
                        uint8_t d, uint32_t* ret) {
  return guarded([&]() -> int {
    if (!t || !ret || op < 0 || op > 2) return setErr(B12X_E_ARG, "bad argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const DepthClass dc = classifyDepth(T, d);
    const Source src = makeSource(T, dIt, check);
    if (dc == DC_BEYOND) {  // no box to hit (:2797)
      *ret = 0;
      return B12X_OK;
    }
    if (dc == DC_GENERAL) {  // :2780-2795 hits become a stencil over the boxes at d
      DepthMap M;
      buildDepthMap(T, d, M);
      runImage(T, *s, src, false, 0, nullptr, &M);
      *ret = depthFlagsOp(T, M, T.hit, op, f, f2);
      hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(std::max<uint32_t>(M.count, 1), kBlock)), dim3(kBlock), 0,
                         T.stream, T.hit, uint64_t(M.count), uint8_t(0));
      HIPCHK(hipStreamSynchronize(T.stream));
      return B12X_OK;
    }
    runImage(T, *s, src, dc == DC_ROOT, 0, nullptr);
    if (dc == DC_ROOT) {  // ImplicitBoxTree.hpp:2722-2748 (surrounding box hit by any valid point)
      uint8_t h0 = 0;
      HIPCHK(hipMemcpyAsync(&h0, T.hit, 1, hipMemcpyDeviceToHost, T.stream));
      hipLaunchKernelGGL(fillU8Kernel, dim3(1), dim3(kBlock), 0, T.stream, T.hit, uint64_t(1), uint8_t(0));
      HIPCHK(hipStreamSynchronize(T.stream));
      const LeafStats ls = leafStats(T);
      if (h0 && rootPred(op, ls, f)) {
        applyFlags(T, nullptr, false, op, f, f2);
        *ret = 1;
      } else {
        *ret = 0;
      }
      return B12X_OK;
    }
    *ret = applyFlags(T, T.hit, true, op, f, f2);
    return B12X_OK;
  });
}

int b12x_search_by_image(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, uint8_t d, uint32_t* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const DepthClass dc = classifyDepth(T, d);
    const Source src = makeSource(T, dIt, check);
    const uint64_t P = src.nBoxes * s->grid.k;
    if (P == 0) return B12X_OK;
    uint32_t* dout = dalloc<uint32_t>(P);
    DepthMap M;
    if (dc == DC_GENERAL) buildDepthMap(T, d, M);
    if (dc == DC_BEYOND)
      hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(P, kBlock)), dim3(kBlock), 0, T.stream, dout, P, kInvalid);
    else
      runImage(T, *s, src, dc == DC_ROOT, 1, dout, dc == DC_GENERAL ? &M : nullptr);
    HIPCHK(hipMemcpyAsync(out, dout, P * 4, hipMemcpyDeviceToHost, T.stream));
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(dout);
    return B12X_OK;
  });
}

int b12x_search_compacted(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, uint8_t d, uint32_t* out,
                          uint64_t* nOut) {
  return guarded([&]() -> int {
    if (!t || !nOut) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const DepthClass dc = classifyDepth(T, d);
    const Source src = makeSource(T, dIt, check);
    if (dc == DC_BEYOND) {
      *nOut = 0;
      return B12X_OK;
    }
    DepthMap M;
    if (dc == DC_GENERAL) buildDepthMap(T, d, M);
    runImage(T, *s, src, dc == DC_ROOT, 0, nullptr, dc == DC_GENERAL ? &M : nullptr);
    const uint32_t m = dc == DC_ROOT ? 1 : (dc == DC_GENERAL ? M.count : T.n);
    if (m == 0) {  // no box at d: nothing can be hit
      *nOut = 0;
      return B12X_OK;
    }
    IndexOp op{};
    op.flags = nullptr;
    op.hit = T.hit;
    op.check = 0;
    op.out = T.list;
    runScan(T, op, m, T.totals, 0);
    syncScal(T, 1);
    const uint32_t cnt = T.hscal[0];
    if (out && cnt) HIPCHK(hipMemcpyAsync(out, T.list, size_t(cnt) * 4, hipMemcpyDeviceToHost, T.stream));
    hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(m, kBlock)), dim3(kBlock), 0, T.stream, T.hit, uint64_t(m),
                       uint8_t(0));
    HIPCHK(hipStreamSynchronize(T.stream));
    *nOut = cnt;
    return B12X_OK;
  });
}

int b12x_search_by_points(b12x_tree t, const double* pts, uint64_t n, uint8_t d, uint32_t* out) {
  return guarded([&]() -> int {
    if (!t || (n && (!pts || !out))) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    const DepthClass dc = classifyDepth(T, d);
    if (n == 0) return B12X_OK;
    double* dp = dalloc<double>(n * T.dim);
    uint32_t* dout = dalloc<uint32_t>(n);
    HIPCHK(hipMemcpyAsync(dp, pts, n * T.dim * 8, hipMemcpyHostToDevice, T.stream));
    const uint32_t nb = ceilDiv(n, kBlock);
    const int d0 = dc == DC_ROOT;
    DepthMap M;
    if (dc == DC_GENERAL) buildDepthMap(T, d, M);
    const uint32_t* mb = dc == DC_GENERAL ? M.box : nullptr;
    const uint8_t* mi = M.in;
    if (dc == DC_BEYOND) {
      hipLaunchKernelGGL(fillU32Kernel, dim3(nb), dim3(kBlock), 0, T.stream, dout, n, kInvalid);
    } else {
      switch (T.dim) {
        case 1: hipLaunchKernelGGL(pointsSearchKernel<1>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, n, T.key[T.cur], T.depth[T.cur], T.n, d0, dout, mb, mi, d); break;
        case 2: hipLaunchKernelGGL(pointsSearchKernel<2>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, n, T.key[T.cur], T.depth[T.cur], T.n, d0, dout, mb, mi, d); break;
        case 3: hipLaunchKernelGGL(pointsSearchKernel<3>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, n, T.key[T.cur], T.depth[T.cur], T.n, d0, dout, mb, mi, d); break;
        default: hipLaunchKernelGGL(pointsSearchKernel<4>, dim3(nb), dim3(kBlock), 0, T.stream, T.dom, dp, n, T.key[T.cur], T.depth[T.cur], T.n, d0, dout, mb, mi, d); break;
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, T.stream));
    HIPCHK(hipStreamSynchronize(T.stream));
    dfree(dp);
    dfree(dout);
    return B12X_OK;
  });
}

int b12x_subdivision_steps(b12x_tree t, b12x_scheme s, uint8_t sd, int nSteps, uint32_t* counts) {
  return guarded([&]() -> int {
    if (!t || nSteps < 0) return setErr(B12X_E_ARG, "bad argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const bool graphs = stepGraphsOn(T);
    // Device scalars: [0] leaves after subdivide, [1] #leaves with sd, [2] leaves kept,
    // [3] #leaves newly hit. One host synchronisation per step.
    if (graphs && nSteps > 0) stageLeafCount(T);
    for (int step = 0; step < nSteps;) {
      const uint32_t n = T.n;
      const uint64_t sub = uint64_t(n) * 2;
      const StepForm form = !graphs ? STEP_EAGER
                            : sub <= kStepSmallLeaves ? STEP_SMALL
                            : sub <= kStepGraphLeaves ? STEP_LARGE
                                                      : STEP_EAGER;
      if (form != STEP_EAGER) {
        // K steps back to back, launched on the stream or replayed from a captured step: the
        // leaf count chains through device scalars and the buffers end in the same half after a
        // step, so the steps need no host input; the leaves can at most double per step, so K
        // keeps every step's subdivided count within the form's bound
        const uint64_t bound = form == STEP_SMALL ? kStepSmallLeaves : kStepGraphLeaves;
        int K = 1;
        while (K < kStepBatch && step + K < nSteps && (uint64_t(n) << (K + 1)) <= bound) ++K;
        ensureCap(T, kStepGraphLeaves);
        const bool viaStream = stepLaunchStream();
        b12x_tree_s::StepGraph& G = T.step[form == STEP_SMALL ? 1 : 0];
        if (!viaStream && (!G.exec || G.scheme != s->id || G.sd != sd || G.gen != T.bufGen || G.cur != T.cur)) {
          G.drop();
          const uint32_t gb = uint32_t(bound / 2);
          HIPCHK(hipStreamBeginCapture(T.stream, hipStreamCaptureModeThreadLocal));
          try {
            if (form == STEP_SMALL) enqueueSmallStep(T, *s, sd);
            else enqueueStep(T, *s, sd, gb, T.scal + 2, form, /*logStep=*/true);
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(T.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          HIPCHK(hipStreamEndCapture(T.stream, &G.graph));
          G.exec = graphReuse() ? ObjectPool::get().takeExec(kGraphKindStep + (form == STEP_SMALL ? 1 : 0)) : nullptr;
          if (G.exec) {  // a released tree's executable: updated in place when the topology matches
            hipGraphNode_t errNode = nullptr;
            hipGraphExecUpdateResult r = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(G.exec, G.graph, &errNode, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
              (void)hipGetLastError();
              (void)hipGraphExecDestroy(G.exec);
              G.exec = nullptr;
            }
          }
          if (!G.exec) HIPCHK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
          G.scheme = s->id;
          G.sd = sd;
          G.gen = T.bufGen;
          G.cur = T.cur;
        }
        if (form == STEP_SMALL) {
          launchKeepSubdivide(T, sd, nullptr, uint32_t(K - 1));  // the first step's subdivide
        } else {
          hipLaunchKernelGGL(fillU32Kernel, dim3(1), dim3(kBlock), 0, T.stream, T.scal + kScalStepLog, uint64_t(1), 0u);
          HIPCHK(hipGetLastError());
        }
        for (int k = 0; k < K; ++k) {
          if (!viaStream) HIPCHK(hipGraphLaunch(G.exec, T.stream));
          else if (form == STEP_SMALL) enqueueSmallStep(T, *s, sd);
          else enqueueStep(T, *s, sd, uint32_t(bound / 2), T.scal + 2, form, /*logStep=*/true);
        }
        HIPCHK(hipStreamSynchronize(T.stream));
        uint32_t kept = 0;
        for (int k = 0; k < K; ++k) {
          const uint32_t* L = T.hscal + kScal + 4 * k;
          const uint32_t nSub = L[0], cSub = L[1], cSet = L[3];
          kept = L[2];
          if (counts) {
            counts[4 * (step + k) + 0] = cSub;
            counts[4 * (step + k) + 1] = cSet;
            counts[4 * (step + k) + 2] = nSub - kept;
            counts[4 * (step + k) + 3] = kept;  // every kept leaf had HIT, which unsetFlags clears
          }
        }
        T.n = kept ? kept : 1;  // the log kernel put the surrounding box in an empty tree
        ++T.leafGen;
        step += K;
      } else {
        ensureCap(T, sub);
        enqueueStep(T, *s, sd, n, nullptr, STEP_EAGER);
        syncScal(T, 4);
        const uint32_t nSub = T.hscal[0], cSub = T.hscal[1], kept = T.hscal[2], cSet = T.hscal[3];
        if (T.timing) {
          T.imagePoints += uint64_t(nSub) * s->grid.k;
          T.launchPoints.push_back(uint64_t(nSub) * s->grid.k);
        }
        T.n = kept;
        ++T.leafGen;
        insertRootIfEmpty(T);
        if (graphs && T.n != kept) stageLeafCount(T);
        if (counts) {
          counts[4 * step + 0] = cSub;
          counts[4 * step + 1] = cSet;
          counts[4 * step + 2] = nSub - kept;
          counts[4 * step + 3] = kept;  // every kept leaf had HIT, which unsetFlags clears
        }
        ++step;
      }
    }
    return B12X_OK;
  });
}

// ------------------------------------------------------------------------------------------------
int b12x_tree_image_timing(b12x_tree t, int enable, double* ms, uint64_t* launches, uint64_t* points) {
  return guarded([&]() -> int {
    if (!t) return setErr(B12X_E_ARG, "null tree");
    HIPCHK(hipStreamSynchronize(t->stream));
    collectTimings(*t);
    if (ms) *ms = t->imageMs;
    if (launches) *launches = t->imageLaunches;
    if (points) *points = t->imagePoints;
    if (enable) {
      t->imageMs = 0.0;
      t->imageLaunches = t->imagePoints = 0;
      t->launchMs.clear();
      t->launchPoints.clear();
    }
    t->timing = enable != 0;
    return B12X_OK;
  });
}

int b12x_tree_image_launches(b12x_tree t, double* ms, uint64_t* points, uint64_t cap, uint64_t* n) {
  return guarded([&]() -> int {
    if (!t || !n) return setErr(B12X_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(t->stream));
    collectTimings(*t);
    const uint64_t k = std::min<uint64_t>(t->launchMs.size(), t->launchPoints.size());
    *n = k;
    for (uint64_t i = 0; i < std::min(k, cap); ++i) {
      if (ms) ms[i] = t->launchMs[i];
      if (points) points[i] = t->launchPoints[i];
    }
    return B12X_OK;
  });
}

int b12x_scheme_create(int dim, int gk, const uint64_t* gp, int ngp, int mk, const double* mp, int nmp,
                       b12x_scheme* out) {
  return guarded([&]() -> int {
    if (!out || dim < 1 || dim > kMaxDim) return setErr(B12X_E_ARG, "bad argument");
    if (int e = ensureDevice()) return e;
    static std::atomic<uint64_t> ids{0};
    auto* S = new b12x_scheme_s();
    S->id = ++ids;
    S->dim = dim;
    S->grid.kind = gk;
    S->map.kind = mk;
    try {
      std::vector<double> gtab;
      buildGrid(dim, gk, gp, ngp, S->grid, gtab);
      if (!gtab.empty()) {
        S->dTable = dalloc<double>(gtab.size());
        HIPCHK(hipMemcpy(S->dTable, gtab.data(), gtab.size() * 8, hipMemcpyHostToDevice));
        S->grid.table = S->dTable;
      }
      std::vector<ForcingStep> tab;
      buildMapParams(dim, mk, mp, nmp, S->map, tab);
      if (mk == MAP_DUFFING) {
        S->dForcing = dalloc<ForcingStep>(tab.size());
        if (!tab.empty())
          HIPCHK(hipMemcpy(S->dForcing, tab.data(), tab.size() * sizeof(ForcingStep), hipMemcpyHostToDevice));
        S->map.forcing = S->dForcing;
      }
    } catch (...) {
      delete S;
      throw;
    }
    *out = S;
    return B12X_OK;
  });
}

int b12x_scheme_destroy(b12x_scheme s) {
  delete s;
  return B12X_OK;
}

int b12x_scheme_points_per_box(b12x_scheme s, uint64_t* out) {
  if (!s || !out) return setErr(B12X_E_ARG, "null argument");
  *out = s->grid.k;
  return B12X_OK;
}

int b12x_map_points(b12x_scheme s, const double* in, double* out, uint64_t n) {
  return guarded([&]() -> int {
    if (!s || (n && (!in || !out))) return setErr(B12X_E_ARG, "null argument");
    if (n == 0) return B12X_OK;
    const int D = s->dim;
    double* di = dalloc<double>(n * D);
    double* dout = dalloc<double>(n * D);
    HIPCHK(hipMemcpy(di, in, n * D * 8, hipMemcpyHostToDevice));
    const uint32_t nb = ceilDiv(n, kBlock);
    switch (s->map.kind) {
      case MAP_IDENTITY:
        if (D == 1) hipLaunchKernelGGL((mapPointsKernel<1, MAP_IDENTITY>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n);
        else if (D == 2) hipLaunchKernelGGL((mapPointsKernel<2, MAP_IDENTITY>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n);
        else if (D == 3) hipLaunchKernelGGL((mapPointsKernel<3, MAP_IDENTITY>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n);
        else hipLaunchKernelGGL((mapPointsKernel<4, MAP_IDENTITY>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n);
        break;
      case MAP_HENON: hipLaunchKernelGGL((mapPointsKernel<2, MAP_HENON>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n); break;
      case MAP_LORENZ: hipLaunchKernelGGL((mapPointsKernel<3, MAP_LORENZ>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n); break;
      case MAP_HENON_HEILES: hipLaunchKernelGGL((mapPointsKernel<4, MAP_HENON_HEILES>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n); break;
      case MAP_DUFFING: hipLaunchKernelGGL((mapPointsKernel<2, MAP_DUFFING>), dim3(nb), dim3(kBlock), 0, 0, s->map, di, dout, n); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dout, n * D * 8, hipMemcpyDeviceToHost));
    dfree(di);
    dfree(dout);
    return B12X_OK;
  });
}

// ------------------------------------------------------------------------------------------------
// Transition matrix
// ------------------------------------------------------------------------------------------------
}  // extern "C"

namespace {

}  // namespace

extern "C" {

int b12x_transition_matrix(b12x_tree t, b12x_scheme s, uint8_t dIt, uint8_t check, uint8_t d, int useAbs,
                           int rowMajor, int useInvalid, b12x_coo* out) {
  return guarded([&]() -> int {
    if (!t || !out) return setErr(B12X_E_ARG, "null argument");
    b12x_tree_s& T = *t;
    checkScheme(T, s);
    const DepthClass dc = classifyDepth(T, d);
    const Source src = makeSource(T, dIt, check);
    const uint32_t k = s->grid.k;
    DepthMap DM;
    if (dc == DC_GENERAL) buildDepthMap(T, d, DM);
    const uint32_t nBoxes =  // count(depth)
        dc == DC_ROOT ? 1u : (dc == DC_GENERAL ? DM.count : (dc == DC_BEYOND ? 0u : T.n));
    auto* M = new b12x_coo_s();
    try {
      streamCreate(&M->stream);
      M->nRows = M->nCols = nBoxes;
      M->rowMajor = rowMajor != 0;
      M->colMajor = rowMajor == 0;
      // passes over whole source boxes, each with at most chunkPoints test points of scratch
      // (ImplicitBoxTree.hpp:2941-3004 sizes its passes from the free memory and carries the open
      // last row across them; a pass here ends at a box boundary, so its rows are complete and
      // the passes' entries follow one another in row-major order)
      const uint64_t bytesPerPoint = k <= 64 ? 4 : 24;  // targets (k > 64: rows, sort buffers)
      const uint64_t budget = tmChunkBudget(T);
      const uint64_t boxesPerPass =
          std::max<uint64_t>(1, std::min<uint64_t>(budget / (bytesPerPoint * k), (uint64_t(1) << 31) / k));
      const uint64_t nRowsSrc = src.nBoxes;
      GrowBuf rows, cols, cnt;
      uint64_t nnz = 0;
      uint32_t passes = 0;
      bool anyInvalid = false;  // scal[2], set by the passes, read with each pass's count
      zeroScal(T);
      for (uint64_t b0 = 0; b0 < nRowsSrc; b0 += boxesPerPass) {
        const uint64_t nb = std::min(boxesPerPass, nRowsSrc - b0), np = nb * k;
        ++passes;
        uint32_t* col = dalloc<uint32_t>(np);
        if (dc != DC_BEYOND)
          runImage(T, *s, src, dc == DC_ROOT, 1, col, dc == DC_GENERAL ? &DM : nullptr, nullptr, b0 * k, np);
        else
          hipLaunchKernelGGL(fillU32Kernel, dim3(ceilDiv(np, kBlock)), dim3(kBlock), 0, T.stream, col, np, kInvalid);
        if (k <= 64) {
          // k targets of a box sorted inside one wave segment, run-length encoded
          uint32_t* rowPtr = dalloc<uint32_t>(nb + 1);
          TmRowArgs a{};
          a.col = col;
          a.nRows = nb;
          a.k = k;
          a.keepInvalid = useInvalid ? 1 : 0;
          a.invalidLabel = nBoxes + 1;  // nColumns_ after its increment (ImplicitBoxTree.hpp:3081-3087,3127-3132)
          a.rowCount = rowPtr;
          a.anyInvalid = T.scal + 2;
          a.pass = 0;
          tmRows(a, T.stream);
          HIPCHK(hipMemsetAsync(rowPtr + nb, 0, 4, T.stream));
          scanArray(T.stream, rowPtr, nb + 1, T.totals, T.scal + 3, nullptr, false);
          syncScal(T, 4);
          const uint32_t nc = T.hscal[3];
          anyInvalid = T.hscal[2] != 0;
          cols.reserve(nnz + nc, nnz, T.stream);
          cnt.reserve(nnz + nc, nnz, T.stream);
          rows.reserve(nnz + nc, nnz, T.stream);
          a.pass = 1;
          a.rowPtr = rowPtr;
          a.outCol = cols.p + nnz;
          a.outCnt = cnt.p + nnz;
          tmRows(a, T.stream);
          hipLaunchKernelGGL(expandRowsKernel, dim3(ceilDiv(nb, kBlock)), dim3(kBlock), 0, T.stream, rowPtr, nb,
                             uint32_t(b0), rows.p + nnz);
          HIPCHK(hipGetLastError());
          dfree(rowPtr);  // the cache fences its reuse behind the queued kernels
          nnz += nc;
        } else {
          // any k: stable sort by column, then by row (row-major), then run-length encoding
          uint32_t* prow = dalloc<uint32_t>(np);
          uint32_t* dummy = dalloc<uint32_t>(np);
          hipLaunchKernelGGL(rowOfPointKernel, dim3(ceilDiv(np, kBlock)), dim3(kBlock), 0, T.stream, prow, np, k,
                             uint32_t(b0));
          hipLaunchKernelGGL(anyInvalidKernel, dim3(ceilDiv(np, kBlock)), dim3(kBlock), 0, T.stream,
                             (const uint32_t*)col, np, T.scal + 2);
          radixSort3(T.stream, col, prow, dummy, np, 32, T.hscal + 4, false);
          radixSort3(T.stream, prow, col, dummy, np, bitsFor(nRowsSrc), T.hscal + 4, false);
          dfree(dummy);
          RunLengthOp op{};
          op.row = prow;
          op.col = col;
          op.n = np;
          op.keepInvalid = useInvalid ? 1 : 0;
          op.invalidLabel = nBoxes + 1;
          // count first to size the outputs (tile totals sized for the pass's points)
          uint32_t* ptot = dalloc<uint32_t>(2 * uint64_t(ceilDiv(np, kTile)) + 16);
          runScan(T, op, uint32_t(np), ptot, 3, nullptr, -1, /*countOnly=*/true);
          syncScal(T, 4);
          const uint32_t nc = T.hscal[3];
          anyInvalid = T.hscal[2] != 0;
          rows.reserve(nnz + nc, nnz, T.stream);
          cols.reserve(nnz + nc, nnz, T.stream);
          cnt.reserve(nnz + nc, nnz, T.stream);
          op.outRow = rows.p + nnz;
          op.outCol = cols.p + nnz;
          op.outCnt = cnt.p + nnz;
          runScan(T, op, uint32_t(np), ptot, 3);
          dfree(ptot);  // the cache fences its reuse behind the queued kernels
          dfree(prow);
          nnz += nc;
        }
        dfree(col);
      }
      lastTmPasses = passes;
      if (nnz > 0xFFFFFFFFull) throw HipFail(B12X_E_UNSUPPORTED, "more than 2^32 - 1 matrix entries");
      if (passes > 1) {  // the doubling's slack
        rows.shrink(nnz, T.stream);
        cols.shrink(nnz, T.stream);
        cnt.shrink(nnz, T.stream);
      }
      if (useInvalid && anyInvalid) M->nCols = nBoxes + 1;
      if (!rowMajor) {
        // column-major order (CooMatrix.hpp:95-137): stable LSD radix sort by column of the
        // row-major entries leaves rows ascending within each column
        radixSort3(T.stream, cols.p, rows.p, cnt.p, nnz, bitsFor(uint64_t(nBoxes) + 1), T.hscal + 4, false);
      }
      M->nnz = uint32_t(nnz);
      M->rows = rows.release();
      M->cols = cols.release();
      M->vals = dalloc<double>(nnz);
      if (nnz)
        hipLaunchKernelGGL(valuesKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, T.stream, (const uint32_t*)cnt.p,
                           uint64_t(nnz), double(k), useAbs ? 1 : 0, M->vals);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(T.stream));
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
    return B12X_OK;
  });
}

int b12x_coo_destroy(b12x_coo m) {
  delete m;
  return B12X_OK;
}

int b12x_coo_info(b12x_coo m, uint64_t* nnz, uint32_t* nr, uint32_t* nc, int* rm) {
  if (!m) return setErr(B12X_E_ARG, "null matrix");
  if (nnz) *nnz = m->nnz;
  if (nr) *nr = m->nRows;
  if (nc) *nc = m->nCols;
  if (rm) *rm = m->rowMajor ? 1 : 0;
  return B12X_OK;
}

int b12x_coo_download(b12x_coo m, uint32_t* rows, uint32_t* cols, double* vals) {
  return guarded([&]() -> int {
    if (!m) return setErr(B12X_E_ARG, "null matrix");
    if (rows && m->nnz) HIPCHK(hipMemcpyAsync(rows, m->rows, m->nnz * 4, hipMemcpyDeviceToHost, m->stream));
    if (cols && m->nnz) HIPCHK(hipMemcpyAsync(cols, m->cols, m->nnz * 4, hipMemcpyDeviceToHost, m->stream));
    if (vals && m->nnz) HIPCHK(hipMemcpyAsync(vals, m->vals, m->nnz * 8, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return B12X_OK;
  });
}

int b12x_coo_upload(const uint32_t* rows, const uint32_t* cols, const double* vals, uint64_t nnz, uint32_t nr,
                    uint32_t nc, b12x_coo* out) {
  return guarded([&]() -> int {
    if (!out || (nnz && (!rows || !cols || !vals))) return setErr(B12X_E_ARG, "null argument");
    for (uint64_t e = 0; e < nnz; ++e)
      if (rows[e] >= nr || cols[e] >= nc) return setErr(B12X_E_ARG, "entry index outside the matrix");
    if (int e = ensureDevice()) return e;
    auto* M = new b12x_coo_s();
    try {
      streamCreate(&M->stream);
      M->nRows = nr;
      M->nCols = nc;
      M->nnz = nnz;
      M->rows = dalloc<uint32_t>(nnz);
      M->cols = dalloc<uint32_t>(nnz);
      M->vals = dalloc<double>(nnz);
      if (nnz) {
        HIPCHK(hipMemcpy(M->rows, rows, nnz * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M->cols, cols, nnz * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M->vals, vals, nnz * 8, hipMemcpyHostToDevice));
      }
      bool rm = true, cm = true;
      for (uint64_t e = 1; e < nnz && rm; ++e)
        rm = rows[e - 1] < rows[e] || (rows[e - 1] == rows[e] && cols[e - 1] < cols[e]);
      for (uint64_t e = 1; e < nnz && cm; ++e)
        cm = cols[e - 1] < cols[e] || (cols[e - 1] == cols[e] && rows[e - 1] < rows[e]);
      M->rowMajor = rm;
      M->colMajor = cm;
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
    return B12X_OK;
  });
}

int b12x_compress_index_vector(const uint32_t* idx, uint64_t n, uint32_t dim, uint32_t* out) {
  return guarded([&]() -> int {
    if ((n && !idx) || !out) return setErr(B12X_E_ARG, "null argument");
    if (n > 0xFFFFFFFFull) return setErr(B12X_E_ARG, "index vector longer than the NrBoxes range");
    if (int e = ensureDevice()) return e;
    uint32_t* di = dalloc<uint32_t>(n);
    uint32_t* dout = dalloc<uint32_t>(uint64_t(dim) + 1);
    if (n) HIPCHK(hipMemcpy(di, idx, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(dim) + 1, kBlock)), dim3(kBlock), 0, 0, di, n, dim, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dout, (uint64_t(dim) + 1) * 4, hipMemcpyDeviceToHost));
    dfree(di);
    dfree(dout);
    return B12X_OK;
  });
}

}  // extern "C"

namespace {
__global__ void sequenceKernel(uint32_t* p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = uint32_t(i);
}

__global__ void gatherF64Kernel(const double* src, const uint32_t* idx, uint64_t n, double* dst) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void gatherU32Kernel(const uint32_t* src, const uint32_t* idx, uint64_t n, uint32_t* dst) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// entries out of order for (major, minor) ascending (equal neighbours are in order)
__global__ void cooOrderKernel(const uint32_t* major, const uint32_t* minor, uint64_t n, uint32_t* bad) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  const bool b = i > 0 && i < n && (major[i] < major[i - 1] || (major[i] == major[i - 1] && minor[i] < minor[i - 1]));
  setFlagOncePerWave(b, bad);
}

// strictly ascending (no repeated entry): the order the matrix flags record
__global__ void cooStrictKernel(const uint32_t* major, const uint32_t* minor, uint64_t n, uint32_t* bad) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  const bool b = i > 0 && i < n && !(major[i] > major[i - 1] || (major[i] == major[i - 1] && minor[i] > minor[i - 1]));
  setFlagOncePerWave(b, bad);
}

bool cooStrict(const b12x_coo_s& M, bool rowMajor) {
  if (M.nnz < 2) return true;
  uint32_t* bad = dalloc<uint32_t>(1);
  HIPCHK(hipMemsetAsync(bad, 0, 4, M.stream));
  hipLaunchKernelGGL(cooStrictKernel, dim3(ceilDiv(M.nnz, kBlock)), dim3(kBlock), 0, M.stream,
                     rowMajor ? M.rows : M.cols, rowMajor ? M.cols : M.rows, M.nnz, bad);
  HIPCHK(hipGetLastError());
  uint32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, M.stream));
  HIPCHK(hipStreamSynchronize(M.stream));
  dfree(bad);
  return h == 0;
}

bool cooSorted(const b12x_coo_s& M, bool rowMajor) {
  if (M.nnz < 2) return true;
  uint32_t* bad = dalloc<uint32_t>(1);
  HIPCHK(hipMemsetAsync(bad, 0, 4, M.stream));
  hipLaunchKernelGGL(cooOrderKernel, dim3(ceilDiv(M.nnz, kBlock)), dim3(kBlock), 0, M.stream,
                     rowMajor ? M.rows : M.cols, rowMajor ? M.cols : M.rows, M.nnz, bad);
  HIPCHK(hipGetLastError());
  uint32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, M.stream));
  HIPCHK(hipStreamSynchronize(M.stream));
  dfree(bad);
  return h == 0;
}
}  // namespace

extern "C" {

// CooMatrix::isSortedRowMajor / isSortedColumnMajor (CooMatrix.hpp:354-374)
int b12x_coo_is_sorted(b12x_coo m, int rowMajor, int* out) {
  return guarded([&]() -> int {
    if (!m || !out) return setErr(B12X_E_ARG, "null argument");
    *out = cooSorted(*m, rowMajor != 0) ? 1 : 0;
    return B12X_OK;
  });
}

// CooMatrix::sortRowMajor / sortColumnMajor (CooMatrix.hpp:95-181, 428-447): stable LSD radix
// sorts by the minor index, then by the major index, carrying the entries' positions
int b12x_coo_sort(b12x_coo m, int rowMajor) {
  return guarded([&]() -> int {
    if (!m) return setErr(B12X_E_ARG, "null matrix");
    b12x_coo_s& M = *m;
    const bool rm = rowMajor != 0;
    const uint64_t nnz = M.nnz;
    if (nnz > 1 && !cooSorted(M, rm)) {
      hipStream_t st = M.stream;
      std::vector<uint32_t> hT(8);
      const uint32_t* major = rm ? M.rows : M.cols;
      const uint32_t* minor = rm ? M.cols : M.rows;
      const uint32_t nMajor = rm ? M.nRows : M.nCols, nMinor = rm ? M.nCols : M.nRows;
      uint32_t* idx = dalloc<uint32_t>(nnz);
      uint32_t* key = dalloc<uint32_t>(nnz);
      uint32_t* dummy = dalloc<uint32_t>(nnz);
      hipLaunchKernelGGL(sequenceKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, idx, nnz);
      HIPCHK(hipMemcpyAsync(key, minor, nnz * 4, hipMemcpyDeviceToDevice, st));
      radixSort3(st, key, idx, dummy, nnz, bitsFor(uint64_t(nMinor) + 1), hT.data());
      hipLaunchKernelGGL(gatherU32Kernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, major,
                         (const uint32_t*)idx, nnz, key);
      radixSort3(st, key, idx, dummy, nnz, bitsFor(uint64_t(nMajor) + 1), hT.data());
      uint32_t* r2 = dalloc<uint32_t>(nnz);
      uint32_t* c2 = dalloc<uint32_t>(nnz);
      double* v2 = dalloc<double>(nnz);
      hipLaunchKernelGGL(gatherU32Kernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, M.rows,
                         (const uint32_t*)idx, nnz, r2);
      hipLaunchKernelGGL(gatherU32Kernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, M.cols,
                         (const uint32_t*)idx, nnz, c2);
      hipLaunchKernelGGL(gatherF64Kernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, M.vals,
                         (const uint32_t*)idx, nnz, v2);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
      dfree(M.rows);
      dfree(M.cols);
      dfree(M.vals);
      M.rows = r2;
      M.cols = c2;
      M.vals = v2;
      dfree(idx);
      dfree(key);
      dfree(dummy);
    }
    if (nnz > 1) {  // the flags record strict orders (no repeated entry)
      M.rowMajor = rm && cooStrict(M, true);
      M.colMajor = !rm && cooStrict(M, false);
    }
    return B12X_OK;
  });
}

}  // extern "C"

namespace {
// the transposed copy of the entries behind them (values negated for the antisymmetric part)
__global__ void cooMirrorKernel(uint32_t* rows, uint32_t* cols, double* vals, uint64_t nnz, int negate) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nnz) return;
  rows[nnz + e] = cols[e];
  cols[nnz + e] = rows[e];
  vals[nnz + e] = negate ? -vals[e] : vals[e];
}

// runs of equal (row, column): a run's first entry holds the run's sum, added from its last entry
// backwards (the reverse inclusive scan of CooMatrix.hpp:405-412), then halved; flag[e] marks run
// starts
__global__ void cooRunSumKernel(const uint32_t* rows, const uint32_t* cols, const double* vals, uint64_t n,
                                double* sum, uint32_t* flag) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= n) return;
  const bool start = e == 0 || rows[e] != rows[e - 1] || cols[e] != cols[e - 1];
  flag[e] = start ? 1u : 0u;
  if (!start) return;
  uint64_t last = e;
  while (last + 1 < n && rows[last + 1] == rows[e] && cols[last + 1] == cols[e]) ++last;
  double acc = vals[last];
  for (uint64_t k = last; k > e; --k) acc = acc + vals[k - 1];
  sum[e] = acc / 2.0;
}

__global__ void cooRunWriteKernel(const uint32_t* rows, const uint32_t* cols, const double* sum, const uint32_t* flag,
                                  const uint32_t* pos, uint64_t n, uint32_t* r2, uint32_t* c2, double* v2) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= n) return;
  const bool start = e == 0 || rows[e] != rows[e - 1] || cols[e] != cols[e - 1];
  if (!start) return;
  const uint32_t o = pos[e];
  r2[o] = rows[e];
  c2[o] = cols[e];
  v2[o] = sum[e];
  (void)flag;
}
}  // namespace

extern "C" {

// CooMatrix::makeSymmetric(makeAntisymmetric) (CooMatrix.hpp:376-426): (P + P^T) / 2 or
// (P - P^T) / 2, in row-major order when P was sorted row-major, column-major otherwise
int b12x_coo_make_symmetric(b12x_coo m, int antisymmetric) {
  return guarded([&]() -> int {
    if (!m) return setErr(B12X_E_ARG, "null matrix");
    b12x_coo_s& M = *m;
    if (M.nRows != M.nCols) return setErr(B12X_E_ARG, "matrix must be square");
    const uint64_t nnz = M.nnz;
    if (nnz == 0) return B12X_OK;
    if (2 * nnz > 0xFFFFFFFFull) return setErr(B12X_E_UNSUPPORTED, "more than 2^31 entries");
    hipStream_t st = M.stream;
    const bool wasRowMajor = cooSorted(M, true);
    const uint64_t n2 = 2 * nnz;
    uint32_t* r = dalloc<uint32_t>(n2);
    uint32_t* c = dalloc<uint32_t>(n2);
    double* v = dalloc<double>(n2);
    HIPCHK(hipMemcpyAsync(r, M.rows, nnz * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c, M.cols, nnz * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(v, M.vals, nnz * 8, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(cooMirrorKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, r, c, v, nnz,
                       antisymmetric ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    dfree(M.rows);
    dfree(M.cols);
    dfree(M.vals);
    M.rows = r;
    M.cols = c;
    M.vals = v;
    M.nnz = n2;
    M.rowMajor = M.colMajor = false;
    if (int e = b12x_coo_sort(m, wasRowMajor ? 1 : 0)) return e;
    DevScope ds;
    double* sum = ds.alloc<double>(n2);
    uint32_t* pos = ds.alloc<uint32_t>(n2);
    uint32_t* tt = ds.alloc<uint32_t>(uint64_t(ceilDiv(n2, kTile)) + 16);
    uint32_t* dTot = ds.alloc<uint32_t>(1);
    uint32_t hTot = 0;
    hipLaunchKernelGGL(cooRunSumKernel, dim3(ceilDiv(n2, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)M.rows,
                       (const uint32_t*)M.cols, (const double*)M.vals, n2, sum, pos);
    HIPCHK(hipGetLastError());
    const uint32_t k = scanArray(st, pos, n2, tt, dTot, &hTot);
    uint32_t* r2 = dalloc<uint32_t>(k);
    uint32_t* c2 = dalloc<uint32_t>(k);
    double* v2 = dalloc<double>(k);
    hipLaunchKernelGGL(cooRunWriteKernel, dim3(ceilDiv(n2, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)M.rows,
                       (const uint32_t*)M.cols, (const double*)sum, (const uint32_t*)nullptr, (const uint32_t*)pos,
                       n2, r2, c2, v2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    for (void* q : {(void*)M.rows, (void*)M.cols, (void*)M.vals}) dfree(q);
    M.rows = r2;
    M.cols = c2;
    M.vals = v2;
    M.nnz = k;
    M.rowMajor = wasRowMajor;  // runs merged: strictly ordered now
    M.colMajor = !wasRowMajor;
    return B12X_OK;
  });
}

// CooMatrix copy constructor (CooMatrix.hpp:58-76): a device copy with its own stream
int b12x_coo_copy(b12x_coo m, b12x_coo* out) {
  return guarded([&]() -> int {
    if (!m || !out) return setErr(B12X_E_ARG, "null argument");
    auto* C = new b12x_coo_s();
    try {
      streamCreate(&C->stream);
      C->nRows = m->nRows;
      C->nCols = m->nCols;
      C->nnz = m->nnz;
      C->rowMajor = m->rowMajor;
      C->colMajor = m->colMajor;
      C->rows = dalloc<uint32_t>(m->nnz);
      C->cols = dalloc<uint32_t>(m->nnz);
      C->vals = dalloc<double>(m->nnz);
      HIPCHK(hipStreamSynchronize(m->stream));
      if (m->nnz) {
        HIPCHK(hipMemcpyAsync(C->rows, m->rows, m->nnz * 4, hipMemcpyDeviceToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->cols, m->cols, m->nnz * 4, hipMemcpyDeviceToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->vals, m->vals, m->nnz * 8, hipMemcpyDeviceToDevice, C->stream));
      }
      HIPCHK(hipStreamSynchronize(C->stream));
    } catch (...) {
      delete C;
      throw;
    }
    *out = C;
    return B12X_OK;
  });
}

// ------------------------------------------------------------------------------------------------
// Invariant measure
// ------------------------------------------------------------------------------------------------
}  // extern "C"

namespace {

// CSC over the first n columns, rows ascending within each column (a column-major copy of the
// entries); entries in column n (the useInvalid label) fall past colPtr[n].
struct Csc {
  uint32_t* keys = nullptr;  // column of each entry (sorted)
  uint32_t* rows = nullptr;
  double* vals = nullptr;
  uint32_t* colPtr = nullptr;  // n + 1
};

// The caller owns the arrays (adoptCsc hands them to a DevScope). They are allocated here
// directly: the radix sorts swap their key and payload pointers with buffers of their own and
// free the originals, so only the final pointers may be owned by anyone.
Csc buildCsc(b12x_coo_s& M, uint32_t n) {
  hipStream_t st = M.stream;
  std::vector<uint32_t> hT(8);
  const uint64_t nnz = M.nnz;
  Csc C;
  C.keys = dalloc<uint32_t>(nnz);
  C.rows = dalloc<uint32_t>(nnz);
  C.vals = dalloc<double>(nnz);
  if (nnz) {
    HIPCHK(hipMemcpyAsync(C.keys, M.cols, nnz * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(C.rows, M.rows, nnz * 4, hipMemcpyDeviceToDevice, st));
  }
  if (M.colMajor) {
    if (nnz) HIPCHK(hipMemcpyAsync(C.vals, M.vals, nnz * 8, hipMemcpyDeviceToDevice, st));
  } else if (nnz) {
    // stable sort by column carrying (row, entry index), after a stable sort by row when the
    // entries are in no order; then gather the values
    uint32_t* idx = dalloc<uint32_t>(nnz);
    hipLaunchKernelGGL(sequenceKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, idx, nnz);
    if (!M.rowMajor) radixSort3(st, C.rows, C.keys, idx, nnz, bitsFor(uint64_t(M.nRows) + 1), hT.data());
    radixSort3(st, C.keys, C.rows, idx, nnz, bitsFor(uint64_t(M.nCols) + 1), hT.data());
    hipLaunchKernelGGL(gatherF64Kernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, M.vals, idx, nnz, C.vals);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    dfree(idx);
  }
  C.colPtr = dalloc<uint32_t>(uint64_t(n) + 1);
  hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(n) + 1, kBlock)), dim3(kBlock), 0, st, C.keys, nnz, n,
                     C.colPtr);
  HIPCHK(hipGetLastError());
  return C;
}

Csc adoptCsc(Csc C, DevScope& ds) {
  ds.adopt(C.keys);
  ds.adopt(C.rows);
  ds.adopt(C.vals);
  ds.adopt(C.colPtr);
  return C;
}

// The CSC of a matrix for reading only: a column-major matrix already holds the rows and values
// in CSC order, so only the column offsets are made (owned by ds); other matrices are sorted into
// a copy (buildCsc).
Csc readCsc(b12x_coo_s& M, uint32_t n, DevScope& ds) {
  if (!M.colMajor) return adoptCsc(buildCsc(M, n), ds);
  Csc C;
  C.keys = M.cols;
  C.rows = M.rows;
  C.vals = M.vals;
  C.colPtr = ds.alloc<uint32_t>(uint64_t(n) + 1);
  hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(n) + 1, kBlock)), dim3(kBlock), 0, M.stream,
                     (const uint32_t*)M.cols, uint64_t(M.nnz), n, C.colPtr);
  HIPCHK(hipGetLastError());
  return C;
}

// Subspace iteration (the oracle's definition; every dot product in dotPartialKernel's strided
// order, which the oracle reproduces): X_0 = cholQR2(start); for k = 0, 1, ...: Y = X_k P, the
// Ritz matrix H = X_k^T Y and, when a tolerance is set or k = maxIter, the residual
// rho = ||Y - X_k H||_F / ||H||_F (the two sums on the device, the square roots and the quotient
// on the host); stop at k = maxIter or rho <= tol, else X_{k+1} = cholQR2(Y).
template <int M>
void subspaceRun(b12x_coo_s& mat, uint32_t n, int lead, int maxIter, double tol, double* ritz, double* basis,
                 int* itersOut, double* residOut) {
  hipStream_t st = mat.stream;
  DevScope ds;
  Csc C = adoptCsc(buildCsc(mat, n), ds);
  const uint32_t nb = ceilDiv(n, kBlock);
  const uint32_t nd = std::min<uint32_t>(nb, kDotBlocks);
  double* X = ds.alloc<double>(uint64_t(M) * n);
  double* Y = ds.alloc<double>(uint64_t(M) * n);
  double* part = ds.alloc<double>(uint64_t(M) * M * nd);
  double* R = ds.alloc<double>(uint64_t(M) * M);
  double* H = ds.alloc<double>(uint64_t(M) * M + 1);  // H, then the residual's sum of squares
  int* bad = ds.alloc<int>(1);
  hipLaunchKernelGGL(fillU32Kernel, dim3(1), dim3(kBlock), 0, st, (uint32_t*)bad, uint64_t(1), 0u);
  auto cholQR = [&](double* Z) {  // in place
    hipLaunchKernelGGL((dotPartialKernel<M, 1>), dim3(nd), dim3(kBlock), 0, st, (const double*)Z, (const double*)Z, n,
                       part);
    hipLaunchKernelGGL(cholKernel<M>, dim3(1), dim3(kBlock), 0, st, (const double*)part, nd, R, bad);
    hipLaunchKernelGGL(applyRinvKernel<M>, dim3(nb), dim3(kBlock), 0, st, Z, n, (const double*)R);
  };
  auto spmm = [&](const double* in, double* out) {
    hipLaunchKernelGGL(spmmKernel<M>, dim3(nb), dim3(kBlock), 0, st, (const uint32_t*)C.colPtr,
                       (const uint32_t*)C.rows, (const double*)C.vals, in, n, out);
  };
  hipLaunchKernelGGL(subspaceStartKernel, dim3(nb), dim3(kBlock), 0, st, X, n, M);
  cholQR(X);
  cholQR(X);
  double hh[M * M + 1];
  int hbad = 0;
  for (int k = 0;; ++k) {
    spmm(X, Y);
    hipLaunchKernelGGL((dotPartialKernel<M, 0>), dim3(nd), dim3(kBlock), 0, st, (const double*)X, (const double*)Y,
                       n, part);
    hipLaunchKernelGGL(ritzKernel, dim3(1), dim3(kBlock), 0, st, (const double*)part, nd, M, H);
    if (tol >= 0.0 || k == maxIter) {
      hipLaunchKernelGGL(residualPartialKernel<M>, dim3(nd), dim3(kBlock), 0, st, (const double*)X,
                         (const double*)Y, (const double*)H, n, lead, part);
      hipLaunchKernelGGL(sumPartialsKernel, dim3(1), dim3(kBlock), 0, st, (const double*)part, nd, H + M * M);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hh, H, sizeof(hh), hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if (hbad) throw HipFail(B12X_E_ARG, "subspace iteration: Gram matrix not positive definite");
      double h2 = 0.0;  // ||H_LL||_F^2, row by row
      for (int a = 0; a < lead; ++a)
        for (int b = 0; b < lead; ++b) h2 += hh[a * M + b] * hh[a * M + b];
      const double rho = std::sqrt(hh[M * M]) / std::sqrt(h2);
      if (k == maxIter || rho <= tol) {
        std::memcpy(ritz, hh, sizeof(double) * M * M);
        if (itersOut) *itersOut = k;
        if (residOut) *residOut = rho;
        break;
      }
    }
    cholQR(Y);
    cholQR(Y);
    std::swap(X, Y);
  }
  if (basis) {
    hipLaunchKernelGGL(transposeBasisKernel, dim3(nb), dim3(kBlock), 0, st, (const double*)X, n, M, Y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(basis, Y, uint64_t(M) * n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
}

// Worker blocks for the persistent power iteration over n columns: one wave per group of 64
// columns up to kFlowMaxWorkers blocks, never more than the device holds resident at once with
// the aggregator block beside them, so that on an idle device every block runs from the start.
// 0: use the batched form (no columns, or n >= 2^29, beyond the 32-bit byte offsets of the
// cached entries).
thread_local double lastPiKernelMs = 0.0;  // duration of the calling thread's last persistent launch
thread_local int lastPiForm = -1;          // its form: 0 straight-line, 1 sweeping; -1 batched alone

uint32_t piFlowWorkers(uint32_t n) {
  if (n == 0 || n >= (1u << 29)) return 0;
  int dev = 0, cus = 0, occ = 0, occ2 = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(&piFlowKernel<true>), kBlock,
                                                   0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, reinterpret_cast<const void*>(&piFlowKernel<false>), kBlock,
                                                   0) != hipSuccess)
    return 0;
  occ = std::min(occ, occ2);
  if (occ <= 0) return 0;
  const uint64_t resident = uint64_t(occ) * uint64_t(cus);
  if (resident < 2) return 0;
  const uint64_t want = ceilDiv(uint64_t(n), kBlock);
  return uint32_t(std::min<uint64_t>({want, uint64_t(kFlowMaxWorkers), resident - 1}));
}

// Device buffers of one solve (sizes for n columns; the persistent form's only when workers > 0).
struct PiWork {
  double* mu0 = nullptr;      // n
  double* ybuf = nullptr;     // 3 n: the batched form's z slots
  double* partial = nullptr;  // 2 ceil(n / kBlock): the batched form's chunk partials
  PiState* pst = nullptr;
  uint64_t* ring = nullptr;   // kFlowRing n: the persistent form's z slots
  uint64_t* part = nullptr;   // kFlowPart ceil(n / 64) 2: its group partials
  uint64_t* tot = nullptr;    // kFlowTotWords: its totals
  int* err = nullptr;
  uint32_t* maxLen = nullptr; // 1: the longest column
  double* dmu = nullptr;      // n: the result
};

// The solve on a CSC matrix (colPtr n + 1, rows ascending within each column): the persistent
// form when workers > 0 and B12X_PI_PERSISTENT is not '0', with the batched form as the fallback
// when a bounded wait gave up (B12X_PI_PERSISTENT=2: an error instead); the batched form alone
// otherwise, its iterations launched in batches from a captured HIP graph unless useGraph is
// false. mu (host, n doubles) is the start vector when muInit, and receives the measure.
// spinLimit bounds every wait of the persistent form (polls); timed records the persistent
// launch's duration with HIP events. With useGraph and timed false nothing is allocated here.
constexpr size_t kPinnedStageMax = size_t(64) << 20;  // measures larger than this copy directly

void piSolve(hipStream_t st, const uint32_t* colPtr, const uint32_t* rows, const double* vals, uint32_t n,
             int maxIter, double tol, int muInit, double* mu, int* iters, double* resid, const PiWork& w,
             uint32_t workers, bool useGraph, uint32_t spinLimit, bool timed) {
  lastPiKernelMs = 0.0;
  lastPiForm = -1;
  const uint32_t nb = std::max<uint32_t>(1, ceilDiv(n, kBlock));
  PiState init{};
  init.t = 1;
  init.resid = -1.0;
  init.done = (maxIter <= 0 || n == 0) ? 1 : 0;
  if (muInit && n) HIPCHK(hipMemcpyAsync(w.mu0, mu, size_t(n) * 8, hipMemcpyHostToDevice, st));
  const uint32_t mi = uint32_t(std::max(maxIter, 0));
  const char* pp = std::getenv("B12X_PI_PERSISTENT");
  bool done = init.done != 0;
  const bool flow = !done && workers && !(pp && pp[0] == '0');
  const uint32_t nG = ceilDiv(n, kWave);
  const uint64_t nRing = uint64_t(kFlowRing) * n, nPart = uint64_t(kFlowPart) * nG * 2;
  {
    // the state, mu_0 and the persistent form's buffers in one launch
    const PiInitArgs ia{w.pst,  init,  w.mu0, n,     (!muInit && n) ? 1 : 0, flow ? w.ring : nullptr, nRing,
                        w.part, nPart, w.tot, w.err, w.maxLen};
    const uint64_t words = std::max<uint64_t>({uint64_t(n), flow ? nRing : 0, flow ? nPart : 0,
                                               flow ? uint64_t(kFlowTotWords) : 0, 1});
    hipLaunchKernelGGL(piInitKernel, dim3(uint32_t(std::min<uint64_t>(ceilDiv(words, kBlock), 65536))), dim3(kBlock),
                       0, st, ia);
    HIPCHK(hipGetLastError());
  }
  const size_t muBytes = size_t(n) * 8;
  constexpr size_t kEarlyState = 64, kEarlyHead = 256;  // pinned block: error word, state, measure
  static_assert(kEarlyState + sizeof(PiState) <= kEarlyHead, "the state fits before the measure");
  // the one-synchronisation read-back of the persistent form (timed calls): a pooled pinned
  // block, returned when the solve ends (after the stream has drained, also on an error)
  struct PinnedLease {
    hipStream_t st = nullptr;
    char* p = nullptr;
    size_t bytes = 0;
    ~PinnedLease() {
      if (!p) return;
      (void)hipStreamSynchronize(st);
      ObjectPool::get().givePinned(p, bytes);
    }
  } lease;
  lease.st = st;
  char*& early = lease.p;
  size_t& earlyBytes = lease.bytes;
  bool earlyDone = false;  // the measure, state and error word arrived with the launch's synchronisation
  if (flow) {
    // the straight-line form when every wave owns one group and every column fits the cache;
    // the kernel itself checks the columns (the longest one is computed on the device, so no
    // host round trip precedes the launch) and leaves at once if one is longer
    hipLaunchKernelGGL(colMaxLenKernel, dim3(std::min<uint32_t>(nb, 256)), dim3(kBlock), 0, st, colPtr, n, w.maxLen);
    HIPCHK(hipGetLastError());
    bool full = uint64_t(nG) > 4ull * workers;
    PiFlowArgs fa{colPtr, rows, vals, w.mu0, w.ring, w.part, w.tot, w.pst, w.err, n, nG, workers, mi, tol, spinLimit,
                  w.maxLen};
    hipEvent_t e0 = nullptr, e1 = nullptr;  // the launch's duration, for measurement
    if (timed) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
    }
    // The allocating entry point (`timed`) reads the outcome, the state and the measure back in
    // one synchronisation: the measure is formed and copied right behind the launch into a pooled
    // pinned block (error word, state, measure) and kept when no wait gave up; otherwise the
    // paths below recompute it.
    if (timed && n && kEarlyHead + muBytes <= kPinnedStageMax) {
      earlyBytes = size_t(1) << 16;
      while (earlyBytes < kEarlyHead + muBytes) earlyBytes <<= 1;
      early = static_cast<char*>(ObjectPool::get().takePinned(earlyBytes));
    }
    int herr = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
      if (timed) HIPCHK(hipEventRecord(e0, st));
      if (full) hipLaunchKernelGGL(piFlowKernel<true>, dim3(workers + 1), dim3(kBlock), 0, st, fa);
      else hipLaunchKernelGGL(piFlowKernel<false>, dim3(workers + 1), dim3(kBlock), 0, st, fa);
      HIPCHK(hipGetLastError());
      if (timed) HIPCHK(hipEventRecord(e1, st));
      if (early) {
        hipLaunchKernelGGL(piFinishKernel, dim3(nb), dim3(kBlock), 0, st, (const double*)w.ring,
                           (const double*)w.mu0, n, (const PiState*)w.pst, w.dmu);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(early, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(early + kEarlyState, w.pst, sizeof(PiState), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(early + kEarlyHead, w.dmu, muBytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::memcpy(&herr, early, sizeof(int));
      } else {
        HIPCHK(hipMemcpyAsync(&herr, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
      }
      if (herr != kFlowNeedsSweep || full) break;
      // a column longer than the cache: every block left before touching the state; the
      // sweeping form runs from the same start
      full = true;
      hipLaunchKernelGGL(fillU32Kernel, dim3(1), dim3(kBlock), 0, st, (uint32_t*)w.err, uint64_t(1), 0u);
      HIPCHK(hipGetLastError());
    }
    lastPiForm = full ? 1 : 0;
    if (timed) {
      float kms = 0.f;
      HIPCHK(hipEventElapsedTime(&kms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      lastPiKernelMs = kms;
    }
    // a wait past its limit (other work holding the CUs): every wave has left; recompute in
    // batches (B12X_PI_PERSISTENT=2: report it instead)
    if (herr && pp && pp[0] == '2') throw HipFail(B12X_E_HIP, "persistent power iteration: a wait passed its limit");
    if (!herr) {
      done = true;
      earlyDone = early != nullptr;
    } else {
      lastPiKernelMs = 0.0;
      lastPiForm = -1;
      HIPCHK(hipMemcpyAsync(w.pst, &init, sizeof(init), hipMemcpyHostToDevice, st));
    }
  }
  const double* zbase = w.ybuf;
  if (done && !init.done && workers && !(pp && pp[0] == '0')) zbase = reinterpret_cast<const double*>(w.ring);
  if (!done) {
    // batched form: step + reduce per iteration, 32 iterations per launch of a captured graph
    // (B12X_GRAPHS=0: eager launches); the kernels do nothing once the loop has ended
    const int batch = 32;
    const char* ge = std::getenv("B12X_GRAPHS");
    useGraph = useGraph && !(ge && ge[0] == '0');
    auto enqueueBatch = [&]() {
      for (int b = 0; b < batch; ++b) {
        hipLaunchKernelGGL(piStepKernel, dim3(nb), dim3(kBlock), 0, st, colPtr, rows, vals, (const double*)w.mu0,
                           w.ybuf, n, w.partial, w.partial + nb, (const PiState*)w.pst);
        hipLaunchKernelGGL(piReduceKernel, dim3(1), dim3(kWave), 0, st, (const double*)w.partial,
                           (const double*)(w.partial + nb), nb, w.pst, mi, tol);
      }
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (useGraph) {
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueueBatch();
      HIPCHK(hipStreamEndCapture(st, &graph));
      HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    PiState hs{};
    for (int64_t it = 0; it <= int64_t(mi) + 2; it += batch) {
      if (useGraph) HIPCHK(hipGraphLaunch(exec, st));
      else enqueueBatch();
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&hs, w.pst, sizeof(hs), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if (hs.done) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  if (earlyDone) {
    PiState hs{};
    std::memcpy(&hs, early + kEarlyState, sizeof(PiState));
    std::memcpy(mu, early + kEarlyHead, muBytes);
    if (iters) *iters = int(hs.final);
    if (resid) *resid = hs.resid;
    return;
  }
  PiState hs{};
  if (n) hipLaunchKernelGGL(piFinishKernel, dim3(nb), dim3(kBlock), 0, st, zbase, (const double*)w.mu0, n,
                           (const PiState*)w.pst, w.dmu);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&hs, w.pst, sizeof(hs), hipMemcpyDeviceToHost, st));
  // the measure to the caller's (pageable) memory: through a pooled pinned block when the call
  // may allocate (`timed`, the allocating entry point), one DMA instead of a staged copy
  const bool staged = timed && n && muBytes <= kPinnedStageMax;
  void* pin = nullptr;
  size_t pinBytes = 0;
  if (staged) {
    pinBytes = size_t(1) << 16;
    while (pinBytes < muBytes) pinBytes <<= 1;
    pin = ObjectPool::get().takePinned(pinBytes);
    HIPCHK(hipMemcpyAsync(pin, w.dmu, muBytes, hipMemcpyDeviceToHost, st));
  } else if (n) {
    HIPCHK(hipMemcpyAsync(mu, w.dmu, muBytes, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  if (staged) {
    std::memcpy(mu, pin, muBytes);
    ObjectPool::get().givePinned(pin, pinBytes);
  }
  if (iters) *iters = int(hs.final);
  if (resid) *resid = hs.resid;
}

// Spin limit of the persistent form's waits (polls): B12X_PI_SPIN overrides the default, so a
// test can make every wait give up and take the fallback.
uint32_t piSpinLimit() {
  const char* e = std::getenv("B12X_PI_SPIN");
  if (e && e[0]) {
    char* end = nullptr;
    const unsigned long v = std::strtoul(e, &end, 10);
    if (end && *end == 0 && v < (1ul << 31)) return uint32_t(v);
  }
  return 1u << 21;
}

// Scratch of b12x_power_iteration_noalloc: one file-scope device array, which every GPU the
// library runs on holds until the process ends, each at its own address. Host code uses only
// the address hipGetSymbolAddress returns on the device (the array's name is a host-side stand-in).
// It holds the buffers of one solve of at most kNoAllocMaxCols columns.
constexpr uint32_t kNoAllocMaxCols = 1u << 17;
constexpr uint64_t noAllocAlign(uint64_t b) { return (b + 255) / 256 * 256; }
constexpr uint64_t kNoAllocBytes =
    noAllocAlign((uint64_t(kNoAllocMaxCols) + 1) * 4) +                         // colPtr
    3 * noAllocAlign(uint64_t(kNoAllocMaxCols) * 8) +                           // mu0, dmu, partial
    noAllocAlign(uint64_t(kNoAllocMaxCols) * 3 * 8) +                           // ybuf
    noAllocAlign(sizeof(PiState)) +                                             // pst
    noAllocAlign(uint64_t(kFlowRing) * kNoAllocMaxCols * 8) +                   // ring
    noAllocAlign(uint64_t(kFlowPart) * (kNoAllocMaxCols / kWave) * 2 * 8) +    // part
    noAllocAlign(uint64_t(kFlowTotWords) * 8) + 2 * noAllocAlign(4);            // tot, err, maxLen
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kScratchDevices = 16;
std::mutex g_scratchMutex;
void* g_scratchPtr[kScratchDevices] = {};

// the scratch's device address on the current device, taken once per device with that device
// current
uint8_t* scratchOnDevice() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) throw HipFail(B12X_E_UNSUPPORTED, "device index beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipFail(B12X_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchPtr[dev] = p;
  }
  return static_cast<uint8_t*>(g_scratchPtr[dev]);
}

// the buffers of one solve of n <= kNoAllocMaxCols columns, carved from the scratch
PiWork carveNoAlloc(uint8_t* base, uint32_t** colPtr) {
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    uint8_t* p = base + off;
    off += noAllocAlign(bytes);
    return p;
  };
  const uint64_t N = kNoAllocMaxCols;
  PiWork w;
  *colPtr = reinterpret_cast<uint32_t*>(take((N + 1) * 4));
  w.mu0 = reinterpret_cast<double*>(take(N * 8));
  w.dmu = reinterpret_cast<double*>(take(N * 8));
  w.partial = reinterpret_cast<double*>(take(N * 8));
  w.ybuf = reinterpret_cast<double*>(take(N * 3 * 8));
  w.pst = reinterpret_cast<PiState*>(take(sizeof(PiState)));
  w.ring = reinterpret_cast<uint64_t*>(take(uint64_t(kFlowRing) * N * 8));
  w.part = reinterpret_cast<uint64_t*>(take(uint64_t(kFlowPart) * (N / kWave) * 2 * 8));
  w.tot = reinterpret_cast<uint64_t*>(take(uint64_t(kFlowTotWords) * 8));
  w.err = reinterpret_cast<int*>(take(4));
  w.maxLen = reinterpret_cast<uint32_t*>(take(4));
  if (off > kNoAllocBytes) throw HipFail(B12X_E_HIP, "scratch layout exceeds the array");
  return w;
}
}  // namespace

extern "C" {

int b12x_power_iteration_kernel_ms(double* ms) {
  if (!ms) return setErr(B12X_E_ARG, "null argument");
  *ms = lastPiKernelMs;
  return B12X_OK;
}

int b12x_power_iteration_form(int* form) {
  if (!form) return setErr(B12X_E_ARG, "null argument");
  *form = lastPiForm;
  return B12X_OK;
}

int b12x_power_iteration(b12x_coo m, int maxIter, double tol, int muInit, double* mu, int* iters, double* resid) {
  return guarded([&]() -> int {
    lastPiKernelMs = 0.0;
    if (!m || !mu) return setErr(B12X_E_ARG, "null argument");
    if (m->nRows != m->nCols && m->nCols != m->nRows + 1) return setErr(B12X_E_ARG, "matrix must be square");
    const uint32_t n = m->nRows;
    hipStream_t st = m->stream;
    DevScope ds;
    const Csc C = readCsc(*m, n, ds);
    PiWork w;
    w.mu0 = ds.alloc<double>(n);
    w.ybuf = ds.alloc<double>(uint64_t(n) * 3);
    w.partial = ds.alloc<double>(uint64_t(std::max<uint32_t>(1, ceilDiv(n, kBlock))) * 2);
    w.pst = ds.alloc<PiState>(1);
    w.dmu = ds.alloc<double>(n);
    const uint32_t workers = (maxIter > 0) ? piFlowWorkers(n) : 0;
    if (workers) {
      w.ring = ds.alloc<uint64_t>(uint64_t(kFlowRing) * n);
      w.part = ds.alloc<uint64_t>(uint64_t(kFlowPart) * ceilDiv(n, kWave) * 2);
      w.tot = ds.alloc<uint64_t>(kFlowTotWords);
      w.err = ds.alloc<int>(1);
      w.maxLen = ds.alloc<uint32_t>(1);
    }
    piSolve(st, C.colPtr, C.rows, C.vals, n, maxIter, tol, muInit, mu, iters, resid, w, workers, true,
            piSpinLimit(), true);
    return B12X_OK;
  });
}

int b12x_power_iteration_noalloc(b12x_coo m, int maxIter, double tol, int muInit, double* mu, int* iters,
                                 double* resid) {
  return guarded([&]() -> int {
    lastPiKernelMs = 0.0;
    if (!m || !mu) return setErr(B12X_E_ARG, "null argument");
    if (m->nRows != m->nCols && m->nCols != m->nRows + 1) return setErr(B12X_E_ARG, "matrix must be square");
    const uint32_t n = m->nRows;
    if (n > kNoAllocMaxCols)
      return setErr(B12X_E_UNSUPPORTED, "b12x_power_iteration_noalloc: " + std::to_string(n) +
                                            " columns exceed its scratch maximum of " +
                                            std::to_string(kNoAllocMaxCols) + " (b12x_power_iteration has no limit)");
    if (!m->colMajor && m->nnz > 0)
      return setErr(B12X_E_ARG,
                    "b12x_power_iteration_noalloc: the matrix must be sorted column-major (transitionMatrix's "
                    "default order, or b12x_coo_sort(m, 0))");
    hipStream_t st = m->stream;
    uint32_t* colPtr = nullptr;
    const PiWork w = carveNoAlloc(scratchOnDevice(), &colPtr);
    // CSC in place: the column-major entries are the columns' rows in ascending order; entries of
    // the extra invalid column (column index n or above) lie past colPtr[n]
    hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(n) + 1, kBlock)), dim3(kBlock), 0, st,
                       (const uint32_t*)m->cols, m->nnz, n, colPtr);
    HIPCHK(hipGetLastError());
    const uint32_t workers = (maxIter > 0) ? piFlowWorkers(n) : 0;
    piSolve(st, colPtr, m->rows, m->vals, n, maxIter, tol, muInit, mu, iters, resid, w, workers, false,
            piSpinLimit(), false);
    return B12X_OK;
  });
}

int b12x_subspace_iteration(b12x_coo m, int nvec, int nlead, int max_iter, double tol, double* ritz, double* basis,
                            int* iters, double* resid) {
  return guarded([&]() -> int {
    if (!m || !ritz || nvec < 1 || nvec > kSubspaceMax || nlead < 1 || nlead > nvec || max_iter < 0)
      return setErr(B12X_E_ARG, "bad argument");
    if (m->nRows != m->nCols && m->nCols != m->nRows + 1) return setErr(B12X_E_ARG, "matrix must be square");
    const uint32_t n = m->nRows;
    if (n < uint32_t(nvec)) return setErr(B12X_E_ARG, "fewer boxes than vectors");
    switch (nvec) {
      case 1: subspaceRun<1>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 2: subspaceRun<2>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 3: subspaceRun<3>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 4: subspaceRun<4>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 5: subspaceRun<5>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 6: subspaceRun<6>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      case 7: subspaceRun<7>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
      default: subspaceRun<8>(*m, n, nlead, max_iter, tol, ritz, basis, iters, resid); break;
    }
    return B12X_OK;
  });
}

// L(NC) for the basis-pass width of cols columns (1 <= cols <= 64): cols rounded up to 8
#define B12X_BASIS_WIDTH(cols, L) \
  do {                            \
    switch (((cols) + 7) / 8) {   \
      case 1: L(8); break;        \
      case 2: L(16); break;       \
      case 3: L(24); break;       \
      case 4: L(32); break;       \
      case 5: L(40); break;       \
      case 6: L(48); break;       \
      case 7: L(56); break;       \
      default: L(64); break;      \
    }                             \
  } while (0)

// Krylov basis of the Krylov-Schur eigensolver: the matrix's columns < n in CSC form, the basis
// V (ncv + 1 vectors of n, column by column), the next vector w and the dot partials, on a
// stream of its own (the matrix may be destroyed first).
struct b12x_krylov_s {
  uint32_t n = 0;
  int ncv = 0;
  uint32_t nd = 0;  // blocks of the basis passes: min(ceil(n / 256), kDotBlocks)
  Csc C;
  double* V = nullptr;
  double* w = nullptr;
  double* part = nullptr;  // kKrylovMaxCols * nd
  double* h = nullptr;     // kHSize: the two Gram-Schmidt coefficient sets, w . w, ||w||
  double* Q = nullptr;     // kKrylovMaxCols^2 rotation
  hipStream_t stream = nullptr;
  static constexpr int kH2 = kKrylovMaxCols, kNorm2 = 2 * kKrylovMaxCols, kBeta = kNorm2 + 1, kHSize = kBeta + 1;
  ~b12x_krylov_s() {
    streamRetire(stream);
    dfree(C.keys);
    dfree(C.rows);
    dfree(C.vals);
    dfree(C.colPtr);
    dfree(V);
    dfree(w);
    dfree(part);
    dfree(h);
    dfree(Q);
    streamRelease(stream);
  }

  // out[0..cols) = x_c . v over the first cols basis columns (v: n doubles on the device)
  void dots(uint32_t cols, const double* v, double* out) {
#define B12X_DOTS(NC) hipLaunchKernelGGL(colDotsKernel<NC>, dim3(nd), dim3(kBlock), 0, stream, (const double*)V, cols, v, n, part)
    B12X_BASIS_WIDTH(cols, B12X_DOTS);
#undef B12X_DOTS
    hipLaunchKernelGGL(colDotsReduceKernel, dim3(cols), dim3(kBlock), 0, stream, (const double*)part, nd, out);
  }

  // w -= V[:, 0..cols) hc; then out[0..cols) = x_c . w, or out[0] = w . w with norm
  void update(uint32_t cols, const double* hc, double* out, bool norm) {
#define B12X_UPDATE(NC) hipLaunchKernelGGL((colUpdateKernel<NC, false>), dim3(nd), dim3(kBlock), 0, stream, (const double*)V, cols, hc, w, n, part)
#define B12X_UPDATE_NORM(NC) hipLaunchKernelGGL((colUpdateKernel<NC, true>), dim3(nd), dim3(kBlock), 0, stream, (const double*)V, cols, hc, w, n, part)
    if (norm) B12X_BASIS_WIDTH(cols, B12X_UPDATE_NORM);
    else B12X_BASIS_WIDTH(cols, B12X_UPDATE);
#undef B12X_UPDATE
#undef B12X_UPDATE_NORM
    hipLaunchKernelGGL(colDotsReduceKernel, dim3(norm ? 1u : cols), dim3(kBlock), 0, stream, (const double*)part, nd, out);
  }
};

int b12x_krylov_create(b12x_coo m, int ncv, b12x_krylov* out) {
  return guarded([&]() -> int {
    if (!m || !out || ncv < 2 || ncv + 1 > kKrylovMaxCols) return setErr(B12X_E_ARG, "ncv must lie in [2, 63]");
    if (m->nRows != m->nCols && m->nCols != m->nRows + 1) return setErr(B12X_E_ARG, "matrix must be square");
    const uint32_t n = m->nRows;
    if (n < uint32_t(ncv) + 1) return setErr(B12X_E_ARG, "fewer boxes than basis vectors");
    auto K = std::make_unique<b12x_krylov_s>();
    K->n = n;
    K->ncv = ncv;
    K->nd = std::min<uint32_t>(ceilDiv(n, kBlock), kDotBlocks);
    K->C = buildCsc(*m, n);  // on the matrix's stream, which the creation waits for below
    HIPCHK(hipStreamSynchronize(m->stream));
    streamCreate(&K->stream);
    K->V = dalloc<double>(uint64_t(ncv + 1) * n);
    K->w = dalloc<double>(n);
    K->part = dalloc<double>(uint64_t(kKrylovMaxCols) * K->nd);
    K->h = dalloc<double>(b12x_krylov_s::kHSize);
    K->Q = dalloc<double>(uint64_t(kKrylovMaxCols) * kKrylovMaxCols);
    *out = K.release();
    return B12X_OK;
  });
}

int b12x_krylov_destroy(b12x_krylov k) {
  delete k;
  return B12X_OK;
}

int b12x_krylov_start(b12x_krylov k, const double* v0, double* norm) {
  return guarded([&]() -> int {
    if (!k) return setErr(B12X_E_ARG, "null basis");
    b12x_krylov_s& K = *k;
    const uint32_t nb = ceilDiv(K.n, kBlock);
    if (v0) HIPCHK(hipMemcpyAsync(K.w, v0, uint64_t(K.n) * 8, hipMemcpyHostToDevice, K.stream));
    else hipLaunchKernelGGL(krylovStartKernel, dim3(nb), dim3(kBlock), 0, K.stream, K.w, K.n);
    // ||v0|| in the fixed dot order (V's column 0 stands in for v0 while the dot runs)
    HIPCHK(hipMemcpyAsync(K.V, K.w, uint64_t(K.n) * 8, hipMemcpyDeviceToDevice, K.stream));
    K.dots(1, K.w, K.h);
    double hh = 0.0;
    HIPCHK(hipMemcpyAsync(&hh, K.h, 8, hipMemcpyDeviceToHost, K.stream));
    HIPCHK(hipStreamSynchronize(K.stream));
    const double beta = std::sqrt(hh);
    if (!(beta > 0.0)) return setErr(B12X_E_ARG, "zero start vector");
    HIPCHK(hipMemcpyAsync(K.h + K.kBeta, &beta, 8, hipMemcpyHostToDevice, K.stream));
    hipLaunchKernelGGL(scaleIntoKernel, dim3(nb), dim3(kBlock), 0, K.stream, (const double*)K.w, K.n,
                       (const double*)(K.h + K.kBeta), K.V);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(K.stream));
    if (norm) *norm = beta;
    return B12X_OK;
  });
}

int b12x_krylov_expand(b12x_krylov k, int p, double* h, double* beta) {
  return guarded([&]() -> int {
    if (!k || !h || !beta || p < 0 || p >= k->ncv) return setErr(B12X_E_ARG, "p must lie in [0, ncv)");
    b12x_krylov_s& K = *k;
    const uint32_t n = K.n, nb = ceilDiv(n, kBlock), cols = uint32_t(p) + 1;
    // w = v_p P
    hipLaunchKernelGGL(spmmKernel<1>, dim3(nb), dim3(kBlock), 0, K.stream, (const uint32_t*)K.C.colPtr,
                       (const uint32_t*)K.C.rows, (const double*)K.C.vals, (const double*)(K.V + uint64_t(p) * n), n,
                       K.w);
    // classical Gram-Schmidt twice in three passes over V: h1 = V^T w; w -= V h1 with h2 = V^T w
    // in the same pass; w -= V h2 with w . w in the same pass
    double hv[b12x_krylov_s::kHSize];
    K.dots(cols, K.w, K.h);
    K.update(cols, K.h, K.h + K.kH2, false);
    K.update(cols, K.h + K.kH2, K.h + K.kNorm2, true);
    HIPCHK(hipMemcpyAsync(hv, K.h, (K.kNorm2 + 1) * 8, hipMemcpyDeviceToHost, K.stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(K.stream));
    // v_{p+1} = w / ||w||
    const double b = std::sqrt(hv[K.kNorm2]);
    for (uint32_t c = 0; c < cols; ++c) h[c] = hv[c] + hv[K.kH2 + c];
    *beta = b;
    if (b > 0.0) {
      HIPCHK(hipMemcpyAsync(K.h + K.kBeta, &b, 8, hipMemcpyHostToDevice, K.stream));
      hipLaunchKernelGGL(scaleIntoKernel, dim3(nb), dim3(kBlock), 0, K.stream, (const double*)K.w, n,
                         (const double*)(K.h + K.kBeta), K.V + uint64_t(p + 1) * n);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(K.stream));
    }
    return B12X_OK;
  });
}

int b12x_krylov_rotate(b12x_krylov k, int m, int kk, const double* q) {
  return guarded([&]() -> int {
    if (!k || !q || m < 1 || m > k->ncv || kk < 1 || kk > m) return setErr(B12X_E_ARG, "need 1 <= k <= m <= ncv");
    b12x_krylov_s& K = *k;
    const uint32_t n = K.n;
    HIPCHK(hipMemcpyAsync(K.Q, q, uint64_t(m) * kk * 8, hipMemcpyHostToDevice, K.stream));
    const uint32_t nb = std::min<uint32_t>(ceilDiv(n, kBlock), kRotateBlocks);
#define B12X_ROTATE(NC) hipLaunchKernelGGL(krylovRotateKernel<NC>, dim3(nb), dim3(kBlock), 0, K.stream, K.V, n, m, kk, (const double*)K.Q)
    B12X_BASIS_WIDTH(uint32_t(kk), B12X_ROTATE);
#undef B12X_ROTATE
    HIPCHK(hipMemcpyAsync(K.V + uint64_t(kk) * n, K.V + uint64_t(m) * n, uint64_t(n) * 8, hipMemcpyDeviceToDevice,
                          K.stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(K.stream));
    return B12X_OK;
  });
}

int b12x_krylov_vector(b12x_krylov k, int c, double* out) {
  return guarded([&]() -> int {
    if (!k || !out || c < 0 || c > k->ncv) return setErr(B12X_E_ARG, "column out of range");
    HIPCHK(hipMemcpyAsync(out, k->V + uint64_t(c) * k->n, uint64_t(k->n) * 8, hipMemcpyDeviceToHost, k->stream));
    HIPCHK(hipStreamSynchronize(k->stream));
    return B12X_OK;
  });
}

namespace {
thread_local SccStats lastSccStats;  // the calling thread's last strong-components run

// In-lists of a matrix's graph for the strong components: the sources (rows) of the entries into
// each column v < n, taken from the column-major order of the entries (sorted into that order
// when the matrix holds another); entries of columns >= n fall past inPtr[n].
struct InLists {
  const uint32_t* src = nullptr;
  uint32_t* inPtr = nullptr;
  uint32_t* ownRows = nullptr;
  uint32_t* ownKeys = nullptr;
  InLists() = default;
  InLists(const InLists&) = delete;
  InLists& operator=(const InLists&) = delete;
  ~InLists() {
    dfree(inPtr);
    dfree(ownRows);
    dfree(ownKeys);
  }
};

void buildInLists(const b12x_coo_s& M, uint32_t n, hipStream_t st, InLists& L) {
  const uint64_t nnz = M.nnz;
  const uint32_t* keys = M.cols;
  L.src = M.rows;
  if (!M.colMajor && nnz > 1) {
    std::vector<uint32_t> hT(8);
    L.ownKeys = dalloc<uint32_t>(nnz);
    L.ownRows = dalloc<uint32_t>(nnz);
    uint32_t* dummy = dalloc<uint32_t>(nnz);
    HIPCHK(hipMemcpyAsync(L.ownKeys, M.cols, nnz * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(L.ownRows, M.rows, nnz * 4, hipMemcpyDeviceToDevice, st));
    if (!M.rowMajor) radixSort3(st, L.ownRows, L.ownKeys, dummy, nnz, bitsFor(uint64_t(M.nRows) + 1), hT.data());
    radixSort3(st, L.ownKeys, L.ownRows, dummy, nnz, bitsFor(uint64_t(M.nCols) + 1), hT.data());
    dfree(dummy);
    keys = L.ownKeys;
    L.src = L.ownRows;
  }
  L.inPtr = dalloc<uint32_t>(uint64_t(n) + 1);
  hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(n) + 1, kBlock)), dim3(kBlock), 0, st, keys, nnz, n,
                     L.inPtr);
  HIPCHK(hipGetLastError());
}
}  // namespace

int b12x_scc_last_stats(uint32_t* rounds, uint32_t* passes, uint32_t* checks) {
  if (rounds) *rounds = lastSccStats.rounds;
  if (passes) *passes = lastSccStats.passes;
  if (checks) *checks = lastSccStats.checks;
  return B12X_OK;
}

// strong components of the matrix's graph (chainRecurrentSet.m:26-39, graphconncomp)
int b12x_coo_scc(b12x_coo m, uint32_t* labels, uint32_t* nComp, uint32_t* nBig) {
  return guarded([&]() -> int {
    if (!m) return setErr(B12X_E_ARG, "null matrix");
    const uint32_t n = m->nRows;
    uint32_t* lab = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
    InLists L;
    buildInLists(*m, n, m->stream, L);
    const SccStats st = sccLabels(m->stream, n, L.inPtr, L.src, lab, nullptr);
    lastSccStats = st;
    if (labels && n) {
      HIPCHK(hipMemcpyAsync(labels, lab, size_t(n) * 4, hipMemcpyDeviceToHost, m->stream));
      HIPCHK(hipStreamSynchronize(m->stream));
    }
    dfree(lab);
    if (nComp) *nComp = st.components;
    if (nBig) *nBig = st.nontrivial;
    return B12X_OK;
  });
}

// chainRecurrentSet.m:18-44: subdivide, the leaves' matrix, strong components, keep the leaves of
// components with more than one leaf
int b12x_chain_recurrent_steps(b12x_tree t, b12x_scheme s, uint8_t sd, int nSteps, uint32_t* counts) {
  return guarded([&]() -> int {
    if (!t || !s || nSteps < 0) return setErr(B12X_E_ARG, "bad argument");
    b12x_tree_s& T = *t;
    constexpr uint8_t kHitFlag = 0x01;
    for (int i = 0; i < nSteps; ++i) {
      uint32_t r = 0;
      if (int e = b12x_subdivide(t, sd, &r)) return e;
      b12x_coo M = nullptr;
      if (int e = b12x_transition_matrix(t, s, kLeafDepth, 0, kLeafDepth, 0, 0, 0, &M)) return e;
      const uint32_t n = T.n;
      SccStats st;
      uint32_t* lab = nullptr;
      try {
        lab = dalloc<uint32_t>(std::max<uint32_t>(n, 1));
        // the marks of the leaves to keep go to the hit bytes, which the flag kernel clears
        InLists L;  // the leaves' matrix is in column-major order
        buildInLists(*M, n, T.stream, L);
        st = sccLabels(T.stream, n, L.inPtr, L.src, lab, T.hit);
        lastSccStats = st;
      } catch (...) {  // the hit bytes stay zero between operations
        if (n) hipLaunchKernelGGL(fillU8Kernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, T.stream, T.hit, uint64_t(n),
                                  uint8_t(0));
        dfree(lab);
        delete M;
        throw;
      }
      dfree(lab);
      delete M;
      (void)applyFlags(T, T.hit, true, 0, kHitFlag, 0);
      if (int e = b12x_remove(t, kHitFlag, &r)) return e;
      if (int e = b12x_unset_flags(t, kHitFlag, kLeafDepth, &r)) return e;
      if (counts) {
        counts[4 * i + 0] = n;
        counts[4 * i + 1] = st.components;
        counts[4 * i + 2] = st.nontrivial;
        counts[4 * i + 3] = T.n;
      }
    }
    return B12X_OK;
  });
}

}  // extern "C"
