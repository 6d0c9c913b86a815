// HIP kernels of the b12x hot path (gfx950, wave64).
//
// Tree storage in HBM: three structure-of-arrays vectors per tree, sorted by comparable key:
//   key[n]   uint64  comparable key K (Box.hpp:65-72) — the descent compares these directly
//   depth[n] uint8
//   flags[n] uint8
// plus a hit byte per leaf used by the image kernel. Set operations are scan-based: a count
// pass, a scan of per-tile totals, and a scatter pass that recomputes the per-element counts.
#pragma once

#include "b12x_device.hpp"

namespace b12x {

constexpr int kBlock = 256;                   // threads per block for streaming kernels
constexpr int kItems = 8;                     // elements per thread in scan kernels
constexpr int kTile = kBlock * kItems;        // elements per block in scan kernels
constexpr int kWave = 64;

// ---------------------------------------------------------------------------------------------
// Wave / block scans
// ---------------------------------------------------------------------------------------------
__device__ inline uint32_t waveInclusiveScan(uint32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ inline uint32_t waveReduceSum(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Exclusive scan over the block; returns the exclusive prefix of v and the block total.
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t waveSums[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t inc = waveInclusiveScan(v);
  if (lane == kWave - 1) waveSums[wave] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    const uint32_t s = waveSums[w];
    before += (w < wave) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

__device__ inline uint32_t blockReduceSum(uint32_t v) {
  __shared__ uint32_t waveSums[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  v = waveReduceSum(v);
  if (lane == 0) waveSums[wave] = v;
  __syncthreads();
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) tot += waveSums[w];
  __syncthreads();
  return tot;
}

// ---------------------------------------------------------------------------------------------
// Generic three-phase scan: Op::count(i) -> number of outputs of element i (0, 1 or 2),
// Op::emit(i, c, pos) writes them at pos. Block totals -> exclusive offsets -> scatter.
// ---------------------------------------------------------------------------------------------
// Element counts come from `n`, or from *nDev when nDev is set (a count produced on the device
// by the previous kernel; the grid is then sized for an upper bound). Op::stat(i) is summed into
// *statCounter by the count pass; Op::after(i) runs for every element in the scatter pass.
struct OpBase {
  __device__ uint32_t stat(uint32_t) const { return 0; }
  __device__ void after(uint32_t) const {}
};

__device__ inline uint32_t elemCount(uint32_t n, const uint32_t* nDev) { return nDev ? *nDev : n; }

template <class Op>
static __global__ __launch_bounds__(kBlock) void scanCountKernel(Op op, uint32_t n, const uint32_t* nDev,
                                                          uint32_t* blockTotals, uint32_t* blockStats) {
  n = elemCount(n, nDev);
  if (uint64_t(blockIdx.x) * kTile >= n && blockIdx.x > 0) return;
  const uint64_t base = uint64_t(blockIdx.x) * kTile + uint64_t(threadIdx.x) * kItems;
  uint32_t s = 0, st = 0;
#pragma unroll
  for (int t = 0; t < kItems; ++t) {
    const uint64_t i = base + t;
    if (i < n) {
      s += op.count(uint32_t(i));
      if (blockStats) st += op.stat(uint32_t(i));
    }
  }
  s = blockReduceSum(s);
  if (threadIdx.x == 0) blockTotals[blockIdx.x] = s;
  if (blockStats) {
    st = blockReduceSum(st);
    if (threadIdx.x == 0) blockStats[blockIdx.x] = st;
  }
}

// Single-block exclusive scan of the block totals in place (m of them, or ceil(*nDev / kTile));
// writes the grand total to *total and, when blockStats is set, the sum of the per-block
// statistics to *statTotal.
static __global__ __launch_bounds__(kBlock) void scanTotalsKernel(uint32_t* v, uint32_t m, const uint32_t* nDev,
                                                           uint32_t* total, const uint32_t* blockStats,
                                                           uint32_t* statTotal) {
  if (nDev) m = uint32_t((uint64_t(*nDev) + kTile - 1) / kTile);
  if (m == 0) m = 1;
  uint32_t carry = 0, stat = 0;
  for (uint32_t off = 0; off < m; off += kTile) {
    uint32_t loc[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int t = 0; t < kItems; ++t) {
      const uint32_t i = off + threadIdx.x * kItems + t;
      loc[t] = i < m ? v[i] : 0u;
      s += loc[t];
      if (blockStats && i < m) stat += blockStats[i];
    }
    uint32_t tot;
    uint32_t ex = blockExclusiveScan(s, tot);
#pragma unroll
    for (int t = 0; t < kItems; ++t) {
      const uint32_t i = off + threadIdx.x * kItems + t;
      if (i < m) v[i] = carry + ex;
      ex += loc[t];
    }
    carry += tot;
  }
  if (blockStats) stat = blockReduceSum(stat);
  if (threadIdx.x == 0) {
    *total = carry;
    if (blockStats) *statTotal = stat;
  }
}

template <class Op>
static __global__ __launch_bounds__(kBlock) void scanScatterKernel(Op op, uint32_t n, const uint32_t* nDev,
                                                            const uint32_t* blockOffsets) {
  n = elemCount(n, nDev);
  if (uint64_t(blockIdx.x) * kTile >= n) return;
  const uint64_t base = uint64_t(blockIdx.x) * kTile + uint64_t(threadIdx.x) * kItems;
  uint32_t c[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int t = 0; t < kItems; ++t) {
    const uint64_t i = base + t;
    c[t] = i < n ? op.count(uint32_t(i)) : 0u;
    s += c[t];
  }
  uint32_t tot;
  uint32_t pos = blockExclusiveScan(s, tot) + blockOffsets[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kItems; ++t) {
    const uint64_t i = base + t;
    if (i < n) {
      if (c[t]) op.emit(uint32_t(i), c[t], pos);
      op.after(uint32_t(i));
    }
    pos += c[t];
  }
}

// The three passes above in one block, for short arrays (a launch instead of three): tiles of
// kTile elements in order, each scanned in the block and emitted after the carry of the tiles
// before it. Writes the total to *total and the statistic to *statTotal; with hostOut set, thread
// 0 then copies scal[0..3] there (a step's counts, read by the host after the stream
// synchronises).
template <class Op>
static __global__ __launch_bounds__(kBlock) void scanSmallKernel(Op op, uint32_t n, const uint32_t* nDev,
                                                          uint32_t* total, uint32_t* statTotal,
                                                          const uint32_t* scal, uint32_t* hostOut) {
  n = elemCount(n, nDev);
  uint32_t carry = 0, stat = 0;
  for (uint32_t off = 0; off < n; off += kTile) {
    const uint64_t base = uint64_t(off) + uint64_t(threadIdx.x) * kItems;
    uint32_t c[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int t = 0; t < kItems; ++t) {
      const uint64_t i = base + t;
      c[t] = i < n ? op.count(uint32_t(i)) : 0u;
      s += c[t];
      if (statTotal && i < n) stat += op.stat(uint32_t(i));
    }
    uint32_t tot;
    uint32_t pos = blockExclusiveScan(s, tot) + carry;
#pragma unroll
    for (int t = 0; t < kItems; ++t) {
      const uint64_t i = base + t;
      if (i < n) {
        if (c[t]) op.emit(uint32_t(i), c[t], pos);
        op.after(uint32_t(i));
      }
      pos += c[t];
    }
    carry += tot;
  }
  if (statTotal) stat = blockReduceSum(stat);
  if (threadIdx.x == 0) {
    *total = carry;
    if (statTotal) *statTotal = stat;
    if (hostOut) {
#pragma unroll
      for (int q = 0; q < 4; ++q) hostOut[q] = scal[q];
    }
  }
}

struct TreeView {
  uint64_t* key;
  uint8_t* depth;
  uint8_t* flags;
};

// subdivide (ImplicitBoxTree.hpp:2883-2930): every leaf with all `mask` flags and depth < 64 is
// replaced in place by its two halves (IncrementDepthByFunctor, ComputeOtherHalfFunctor). The
// halves of a leaf are adjacent to it in key order, so the output stays sorted without a sort.
struct SubdivideOp : OpBase {
  TreeView in, out;
  uint8_t mask;
  __device__ uint32_t stat(uint32_t i) const { return areFlagsSet(in.flags[i], mask) ? 1u : 0u; }
  __device__ uint32_t count(uint32_t i) const {
    return (areFlagsSet(in.flags[i], mask) && in.depth[i] < kMaxDepth) ? 2u : 1u;
  }
  __device__ void emit(uint32_t i, uint32_t c, uint32_t pos) const {
    const uint64_t k = in.key[i];
    const int d = in.depth[i];
    const uint8_t f = in.flags[i];
    if (c == 1) {
      out.key[pos] = k;
      out.depth[pos] = uint8_t(d);
      out.flags[pos] = f;
    } else {
      const uint64_t b = bitsOf(k, d) << 1;
      out.key[pos] = keyOf(b, d + 1);
      out.key[pos + 1] = keyOf(b | 1ull, d + 1);
      out.depth[pos] = out.depth[pos + 1] = uint8_t(d + 1);
      out.flags[pos] = out.flags[pos + 1] = f;
    }
  }
};

// remove (ImplicitBoxTree.hpp:2343-2371): stable compaction keeping leaves with all `mask` flags.
// With `hit` set, a leaf counts as carrying `hitFlag` when its hit byte is set (the fused
// setFlagsBySearch + remove of one subdivision step), and `clearMask` is unset in the output
// (the step's trailing unsetFlags).
struct KeepOp : OpBase {
  TreeView in, out;
  uint8_t* hit;
  uint8_t mask, hitFlag, clearMask;
  int clearHit;  // zero the hit byte after use (scatter pass)
  // leaves newly gaining hitFlag (setFlagsBySearch's return value)
  __device__ uint32_t stat(uint32_t i) const { return (hit && hit[i] && !areFlagsSet(in.flags[i], hitFlag)) ? 1u : 0u; }
  __device__ void after(uint32_t i) const {
    if (clearHit) hit[i] = 0;
  }
  __device__ uint8_t flagsAt(uint32_t i) const {
    uint8_t f = in.flags[i];
    if (hit && hit[i]) f |= hitFlag;
    return f;
  }
  __device__ uint32_t count(uint32_t i) const { return areFlagsSet(flagsAt(i), mask) ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const {
    out.key[pos] = in.key[i];
    out.depth[pos] = in.depth[i];
    out.flags[pos] = uint8_t(flagsAt(i) & ~clearMask);
  }
};

// unsubdivide (ImplicitBoxTree.hpp:3413-3461): first pass marks and lifts unsubdividable leaves
// (Functors.hpp:337-377 over the predecessor/successor stencil of ImplicitBoxTree.hpp:2023-2058);
// the compaction keeps the first of each run of equal boxes with the AND of the run's flags.
struct UnsubMarkKernelArgs {
  TreeView in;
  uint32_t n;
  uint8_t mask;
  uint64_t* liftedKey;
  uint8_t* liftedDepth;
  uint32_t* nUnsub;
};

__device__ inline bool isUnsubdividableAt(const TreeView& t, uint32_t n, uint32_t i, uint8_t mask) {
  const int d = t.depth[i];
  if (!(areFlagsSet(t.flags[i], mask) && d > 0)) return false;
  const uint64_t bits = bitsOf(t.key[i], d);
  const uint32_t j = (bits & 1ull) ? (i > 0 ? i - 1 : 0) : (i + 1 < n ? i + 1 : n - 1);
  const int dn = t.depth[j];
  const uint64_t bn = bitsOf(t.key[j], dn);
  // parents: decrementDepthBy(., 1)
  const uint64_t pb = bits >> 1;
  const int pd = d - 1;
  const uint64_t nb = dn > 0 ? (bn >> 1) : bn;
  const int nd = dn > 0 ? dn - 1 : 0;
  if (pb == nb && pd == nd) return areFlagsSet(t.flags[j], mask);
  // !hasSamePath(parent, neighbour)
  const int diff = pd - dn;
  return shr(pb, diff) != shr(bn, -diff);
}

static __global__ __launch_bounds__(kBlock) void unsubMarkKernel(UnsubMarkKernelArgs a) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t u = 0;
  if (i < a.n) {
    const bool un = isUnsubdividableAt(a.in, a.n, i, a.mask);
    int d = a.in.depth[i];
    uint64_t k = a.in.key[i];
    if (un) {
      const uint64_t b = bitsOf(k, d) >> 1;
      d -= 1;
      k = keyOf(b, d);
    }
    a.liftedKey[i] = k;
    a.liftedDepth[i] = uint8_t(d);
    u = un ? 1u : 0u;
  }
  u = blockReduceSum(u);
  if (threadIdx.x == 0 && u) atomicAdd(a.nUnsub, u);
}

struct UnsubMergeOp : OpBase {
  const uint64_t* key;
  const uint8_t* depth;
  const uint8_t* flags;
  uint32_t n;
  TreeView out;
  __device__ bool same(uint32_t a, uint32_t b) const { return key[a] == key[b] && depth[a] == depth[b]; }
  __device__ uint32_t count(uint32_t i) const { return (i == 0 || !same(i, i - 1)) ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const {
    uint8_t f = flags[i];
    for (uint32_t j = i + 1; j < n && same(j, i); ++j) f &= flags[j];
    out.key[pos] = key[i];
    out.depth[pos] = depth[i];
    out.flags[pos] = f;
  }
};

// ---------------------------------------------------------------------------------------------
// Flag operations at leaf depth (ImplicitBoxTree.hpp:2803-2881, 3334-3411, 912-1058 and the
// *BySearch forms :2716-2801, 3247-3332, 822-910). op 0 = set, 1 = unset, 2 = change.
// The return value counts leaves whose state the operation changes, as the reference does.
// ---------------------------------------------------------------------------------------------
__device__ inline bool flagPred(int op, uint8_t f, uint8_t a) {
  return op == 0 ? !areFlagsSet(f, a) : (op == 1 ? isAnyFlagSet(f, a) : areFlagsSet(f, a));
}
__device__ inline uint8_t flagApply(int op, uint8_t f, uint8_t a, uint8_t b) {
  return op == 0 ? uint8_t(f | a)
                 : (op == 1 ? uint8_t(f & ~a) : (areFlagsSet(f, a) ? uint8_t((f & ~a) | b) : f));
}

// sel: per-leaf selector bytes (stencil or hit bytes), nullptr = all leaves. With clearSel the
// selector bytes are zeroed after use (hit bytes are reused by the next image step).
static __global__ __launch_bounds__(kBlock) void flagsKernel(uint8_t* flags, uint32_t n, uint8_t* sel, int clearSel, int op,
                                                      uint8_t a, uint8_t b, uint32_t* counter) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (i < n) {
    const bool on = sel ? sel[i] != 0 : true;
    if (on) {
      const uint8_t f = flags[i];
      c = flagPred(op, f, a) ? 1u : 0u;
      flags[i] = flagApply(op, f, a, b);
      if (clearSel) sel[i] = 0;
    }
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) atomicAdd(counter, c);
}

// ---------------------------------------------------------------------------------------------
// The fused image kernel: one lane per test point. Point i of source box b = i / k:
// box geometry -> grid offset -> x = c + g*r (Functors.hpp:476-511) -> map -> key -> descent.
// mode 0: hit[target] = 1; mode 1: out[i] = target.
// ---------------------------------------------------------------------------------------------
struct ImageArgs {
  Domain dom;
  MapParams map;
  GridParams grid;
  const uint64_t* key;      // leaves (descent target)
  const uint8_t* depth;
  uint32_t n;
  const uint64_t* srcKey;   // source boxes
  const uint8_t* srcDepth;
  const uint32_t* srcList;  // optional indirection into srcKey/srcDepth
  uint64_t nPoints;
  const uint32_t* nBoxesDev;  // when set, nPoints = *nBoxesDev * k (grid sized for a bound)
  uint64_t pointBase;       // global index of point 0 (grid index offset)
  int mode;
  int depth0;               // search on depth 0 (root only)
  uint8_t* hit;
  uint32_t* out;
  const uint64_t* samples;  // descent samples of key[0..n) (sampleKeysKernel), or null
  // search among the boxes at a depth between the root and the leaves (null: leaf search)
  const uint32_t* depthBox;  // box index of every leaf
  const uint8_t* depthIn;    // leaf belongs to a box at that depth
  uint8_t searchDepth;
  uint64_t* outKey;          // when set, the key of every image point (insertion)
};

// samples[j] = keys[pos(j)], j < min(kDescentSamples, n); n from nDev when given
static __global__ __launch_bounds__(kBlock) void sampleKeysKernel(const uint64_t* keys, uint32_t n0,
                                                                 const uint32_t* nDev, uint64_t* samples) {
  const uint32_t n = nDev ? *nDev : n0;
  const uint32_t S = min(kDescentSamples, n);
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j < S) samples[j] = keys[samplePos(j, n, S)];
}

template <int DIM, int MAP, int GRID>
__device__ inline void imagePoint(const ImageArgs& a, uint64_t i, uint64_t gi, uint64_t b, uint32_t gl, uint32_t n,
                                  const uint64_t* sm, uint32_t S) {
  const uint32_t src = a.srcList ? a.srcList[b] : uint32_t(b);
  double c[DIM], rb[DIM], g[DIM], x[DIM], y[DIM];
  boxCoords<DIM>(a.dom, a.srcKey[src], a.srcDepth[src], c, rb);
  if constexpr (GRID == GRID_INNER) {
#pragma unroll
    for (int j = 0; j < DIM; ++j) g[j] = a.grid.table[gl * DIM + j];
  } else {
    uniformGridPoint<DIM>(a.grid.seed, gi, g);
  }
#pragma unroll
  for (int j = 0; j < DIM; ++j) x[j] = c[j] + g[j] * rb[j];
  applyMap<DIM, MAP>(a.map, x, y);
  const uint64_t pk = pointKey<DIM>(a.dom, y);
  uint32_t t;
  if (a.depth0) t = (pk == kInvalidKey) ? kInvalid : 0u;
  else if (a.depthBox) t = locateInDepth(a.key, a.depth, n, pk, a.searchDepth, a.depthIn, a.depthBox);
  else if (S) t = locateStaged(a.key, a.depth, n, pk, sm, S);
  else t = locate(a.key, a.depth, n, pk);
  if (a.outKey) a.outKey[i] = pk;
  if (a.mode == 0) {
    if (t != kInvalid) a.hit[t] = 1;
  } else {
    a.out[i] = t;
  }
}

// PPL points of one lane at once (local indices idx[q], idx[0] < np): lanes past the end repeat
// point idx[0] and store nothing. Same arithmetic per point as imagePoint.
template <int DIM, int MAP, int GRID, int PPL>
__device__ inline void imagePointsN(const ImageArgs& a, const uint32_t (&idx)[PPL], uint32_t np, uint32_t n) {
  const uint32_t k = a.grid.k, base = uint32_t(a.pointBase);
  double x[PPL][DIM], y[PPL][DIM];
  bool valid[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    valid[q] = idx[q] < np;
    const uint32_t gi = base + (valid[q] ? idx[q] : idx[0]);
    const uint32_t b = gi / k;
    const uint32_t src = a.srcList ? a.srcList[b] : b;
    double c[DIM], rb[DIM], g[DIM];
    boxCoords<DIM>(a.dom, a.srcKey[src], a.srcDepth[src], c, rb);
    if constexpr (GRID == GRID_INNER) {
      const uint32_t gl = gi - b * k;
#pragma unroll
      for (int j = 0; j < DIM; ++j) g[j] = a.grid.table[gl * DIM + j];
    } else {
      uniformGridPoint<DIM>(a.grid.seed, gi, g);
    }
#pragma unroll
    for (int j = 0; j < DIM; ++j) x[q][j] = c[j] + g[j] * rb[j];
  }
  applyMapN<DIM, MAP, PPL>(a.map, x, y);
  uint64_t pk[PPL];
  uint32_t t[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) pk[q] = pointKey<DIM>(a.dom, y[q]);
  if (a.depth0) {
#pragma unroll
    for (int q = 0; q < PPL; ++q) t[q] = (pk[q] == kInvalidKey) ? kInvalid : 0u;
  } else {
    locateN<PPL>(a.key, a.depth, n, pk, t);
  }
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    if (!valid[q]) continue;
    if (a.mode == 0) {
      if (t[q] != kInvalid) a.hit[t[q]] = 1;
    } else {
      a.out[idx[q]] = t[q];
    }
  }
}

// Any block size up to kBlock (the launch picks it); PPL points per lane.
template <int DIM, int MAP, int GRID, int PPL>
static __global__ __launch_bounds__(kBlock) void imageKernel(ImageArgs a) {
  const uint32_t BS = blockDim.x;
  __shared__ uint64_t sm[PPL == 1 ? kDescentSamples : 1];
  // a device-side count is capped by the host's bound (a prefix when the launch is a sample)
  const uint64_t nPoints = a.nBoxesDev ? min(uint64_t(*a.nBoxesDev) * a.grid.k, a.nPoints) : a.nPoints;
  const uint32_t n = a.nBoxesDev ? *a.nBoxesDev : a.n;
  if (uint64_t(blockIdx.x) * BS * PPL >= nPoints) return;  // grid sized for a bound: no points here
  if constexpr (PPL > 1) {
    if (a.pointBase + nPoints + uint64_t(PPL) * BS <= 0xFFFFFFFFull &&
        uint64_t(gridDim.x) * BS * PPL <= 0xFFFFFFFFull) {
      const uint32_t np = uint32_t(nPoints), stride = gridDim.x * BS * PPL;
      for (uint32_t i0 = blockIdx.x * BS * PPL + threadIdx.x; i0 < np; i0 += stride) {
        uint32_t idx[PPL];
#pragma unroll
        for (int q = 0; q < PPL; ++q) idx[q] = i0 + q * BS;
        imagePointsN<DIM, MAP, GRID, PPL>(a, idx, np, n);
      }
      return;
    }
  }
  // one point per lane (and every launch too large for 32-bit indices)
  // upper levels of the descent in LDS: the samples of the leaf keys when given, or every key when
  // the tree has at most kDescentSamples leaves (the sample positions are then the identity, so
  // the search runs in LDS up to the final containment test)
  uint32_t S = 0;
  if (PPL == 1 && !a.depth0 && !a.depthBox && n > 0) {
    const uint64_t* src = a.samples ? a.samples : (n <= kDescentSamples ? a.key : nullptr);
    if (src) {
      S = min(kDescentSamples, n);
      for (uint32_t j = threadIdx.x; j < S; j += BS) sm[j] = src[j];
      __syncthreads();
    }
  }
  if (PPL == 1 && a.pointBase + nPoints <= 0xFFFFFFFFull && uint64_t(gridDim.x) * BS <= 0xFFFFFFFFull) {
    // 32-bit index arithmetic (the common case)
    const uint32_t k = a.grid.k, base = uint32_t(a.pointBase), np = uint32_t(nPoints);
    const uint32_t stride = gridDim.x * BS;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < np; i += stride) {
      const uint32_t gi = base + i;
      const uint32_t b = gi / k;
      imagePoint<DIM, MAP, GRID>(a, i, gi, b, gi - b * k, n, sm, S);
    }
  } else {
    const uint64_t stride = uint64_t(gridDim.x) * BS * PPL;
    for (uint64_t i0 = uint64_t(blockIdx.x) * BS * PPL + threadIdx.x; i0 < nPoints; i0 += stride)
      for (int q = 0; q < PPL; ++q) {
        const uint64_t i = i0 + uint64_t(q) * BS;
        if (i >= nPoints) break;
        const uint64_t gi = a.pointBase + i;
        const uint64_t b = gi / a.grid.k;
        imagePoint<DIM, MAP, GRID>(a, i, gi, b, uint32_t(gi - b * a.grid.k), n, sm, S);
      }
  }
}

// Keys of arbitrary points (searchByPoints, ImplicitBoxTree.hpp:2588-2620)
template <int DIM>
static __global__ __launch_bounds__(kBlock) void pointsSearchKernel(Domain dom, const double* pts, uint64_t m,
                                                             const uint64_t* key, const uint8_t* depth, uint32_t n,
                                                             int depth0, uint32_t* out, const uint32_t* depthBox,
                                                             const uint8_t* depthIn, uint8_t d,
                                                             uint64_t* outKey = nullptr) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= m) return;
  double p[DIM];
#pragma unroll
  for (int j = 0; j < DIM; ++j) p[j] = pts[i * DIM + j];
  const uint64_t pk = pointKey<DIM>(dom, p);
  if (outKey) outKey[i] = pk;
  if (depth0) out[i] = (pk == kInvalidKey) ? kInvalid : 0u;
  else if (depthBox) out[i] = locateInDepth(key, depth, n, pk, d, depthIn, depthBox);
  else out[i] = locate(key, depth, n, pk);
}

// Boxes at depth d: in[i] (leaf i belongs to one: depth >= d, or unsubdividable at the
// unsubdividable depth), startFlag[i] and start[i] (leaf i is the first leaf of its box; start
// is scanned into the box indices by depthBoxKernel)
static __global__ __launch_bounds__(kBlock) void depthMapKernel(TreeView v, uint32_t n, uint8_t d, uint32_t* start,
                                                               uint8_t* startFlag, uint8_t* in) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint64_t b, bp;
  int td, tdp;
  truncateTo(v.key[i], v.depth[i], d, b, td);
  const uint32_t ip = i > 0 ? i - 1 : 0;
  truncateTo(v.key[ip], v.depth[ip], d, bp, tdp);
  const bool deep = d == kUnsubDepth || td >= int(d);  // hasAtLeastDepth of the truncated box
  const bool fresh = deep && (i == 0 || !(b == bp && td == tdp));
  const bool un = d == kUnsubDepth ? isUnsubdividableAt(v, n, i, 0) : true;
  const bool inBox = d == kUnsubDepth ? un : int(v.depth[i]) >= int(d);
  const bool st = fresh && un;
  in[i] = inBox ? 1 : 0;
  startFlag[i] = st ? 1 : 0;
  start[i] = st ? 1u : 0u;
}

// exclusive scan of the starts -> box index (inclusive count minus one, 0 before the first box)
static __global__ __launch_bounds__(kBlock) void depthBoxKernel(uint32_t* box, const uint8_t* startFlag, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t inc = box[i] + startFlag[i];
  box[i] = inc > 0 ? inc - 1 : 0;
}

// Flags combined over the leaves of every box (AND for set and change, OR for unset)
static __global__ __launch_bounds__(kBlock) void depthCombineKernel(const uint8_t* flags, uint32_t n,
                                                                   const uint32_t* box, const uint8_t* in, int useOr,
                                                                   uint32_t* acc) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || (in && !in[i])) return;
  if (useOr) atomicOr(acc + box[i], uint32_t(flags[i]));
  else atomicAnd(acc + box[i], uint32_t(flags[i]));
}

// Flag operation on every leaf of a selected box at depth d; counts (into *count) the selected
// boxes whose combined flags pass the operation's predicate (ImplicitBoxTree.hpp:1161-1209)
static __global__ __launch_bounds__(kBlock) void depthFlagsKernel(uint8_t* flags, uint32_t n, const uint32_t* box,
                                                                 const uint8_t* in, const uint8_t* startFlag,
                                                                 const uint8_t* sel, const uint32_t* acc, int op,
                                                                 uint8_t a, uint8_t b2, uint32_t* count) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t c = 0;
  if (i < n) {
    const uint32_t bx = box[i];
    if (startFlag[i] && sel[bx]) {
      c = flagPred(op, uint8_t(acc[bx]), a) ? 1u : 0u;
    }
    if (in[i] && sel[bx]) flags[i] = flagApply(op, flags[i], a, b2);
  }
  c = blockReduceSum(c);
  if (threadIdx.x == 0 && c) atomicAdd(count, c);
}

// Sources at depth d: the first leaf of every box whose leaves all carry `check`, truncated to d
static __global__ __launch_bounds__(kBlock) void depthCheckKernel(const uint8_t* flags, uint32_t n, const uint32_t* box,
                                                                 const uint8_t* in, uint8_t check, uint32_t* ok) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !in[i]) return;
  if ((flags[i] & check) != check) atomicAnd(ok + box[i], 0u);
}

// ---------------------------------------------------------------------------------------------
// deleteDepth (ImplicitBoxTree.hpp:1215-1262) and insertion (ImplicitBoxTree.hpp:1613-1878)
// ---------------------------------------------------------------------------------------------
// box (bits, depth) of a comparable key
__device__ inline uint64_t bitsAt(uint64_t key, int depth) { return bitsOf(key, depth); }

// hasSamePath (Box.hpp:74-83) of two valid boxes: one is an ancestor of (or equal to) the other
__device__ inline bool samePath(uint64_t b1, int d1, uint64_t b2, int d2) {
  const int diff = d1 - d2;
  const uint64_t x = diff > 0 ? (diff >= 64 ? 0ull : b1 >> diff) : b1;
  const uint64_t y = diff < 0 ? (-diff >= 64 ? 0ull : b2 >> -diff) : b2;
  return x == y;
}

// computeDepthOfMutualAncestor (Box.hpp:206-220)
__device__ inline int mutualAncestorDepth(uint64_t b1, int d1, uint64_t b2, int d2) {
  const uint64_t x1 = (64 - d1) >= 64 ? 0ull : b1 << (64 - d1);
  const uint64_t x2 = (64 - d2) >= 64 ? 0ull : b2 << (64 - d2);
  const uint64_t x = x1 ^ x2;
  const int lz = x ? __clzll((long long)x) : 64;
  return min(lz, min(d1, d2));
}

// deleteDepth: leaf i truncated to d1 = depth - 1 starts a run when it differs from leaf i - 1
static __global__ __launch_bounds__(kBlock) void truncRunKernel(TreeView v, uint32_t n, uint8_t d1, uint32_t* start,
                                                               uint8_t* startFlag) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint64_t b, bp;
  int td, tdp;
  truncateTo(v.key[i], v.depth[i], d1, b, td);
  const uint32_t ip = i > 0 ? i - 1 : 0;
  truncateTo(v.key[ip], v.depth[ip], d1, bp, tdp);
  const bool st = i == 0 || !(b == bp && td == tdp);
  start[i] = st ? 1u : 0u;
  startFlag[i] = st ? 1 : 0;
}

// the first leaf of each run writes the truncated box with the AND of the run's flags
static __global__ __launch_bounds__(kBlock) void truncWriteKernel(TreeView in, TreeView out, uint32_t n, uint8_t d1,
                                                                 const uint8_t* startFlag, const uint32_t* run,
                                                                 const uint32_t* acc) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !startFlag[i]) return;
  uint64_t b;
  int td;
  truncateTo(in.key[i], in.depth[i], d1, b, td);
  const uint32_t r = run[i];
  out.key[r] = keyOf(b, td);
  out.depth[r] = uint8_t(td);
  out.flags[r] = uint8_t(acc[r]);
}

// insertion: per point, whether it is inserted (ins) and which existing box it hits (hit[t] = 1)
// form 0: leaf depth (a point held by a leaf shallower than 64 is inserted too), 1: a depth with
// boxes (inserted if no box holds it), 2: deeper than every leaf (every valid point)
// Outside leaf depth a new box is the point's box truncated to depth d (shift = 64 - d), so a
// point whose box equals the previous point's, with both inserted, adds nothing to the merge
// (runs on one path reduce with max and OR): insU32, the compaction flag, drops it, while ins
// (isToInsert) still reports it.
static __global__ __launch_bounds__(kBlock) void insertMarkKernel(const uint64_t* pk, const uint32_t* sv, uint64_t P,
                                                                 int form, int shift, const uint8_t* leafDepth,
                                                                 uint8_t* ins, uint32_t* insU32, uint8_t* hit,
                                                                 uint32_t* anyValid) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= P) return;
  const uint64_t p = pk[i];
  const bool valid = p != kInvalidKey;
  const uint32_t t = form == 2 ? kInvalid : sv[i];
  bool in;
  if (form == 0) in = t == kInvalid ? valid : leafDepth[t] < kMaxDepth;
  else in = t == kInvalid && valid;
  bool keep = in;
  if (in && form != 0 && i > 0) {
    const uint64_t q = pk[i - 1];
    const bool prevIn = q != kInvalidKey && (form == 2 || sv[i - 1] == kInvalid);
    if (prevIn && (q >> shift) == (p >> shift)) keep = false;
  }
  ins[i] = in ? 1 : 0;
  insU32[i] = keep ? 1u : 0u;
  if (t != kInvalid && !(form == 0 && in)) hit[t] = 1;
  if (valid && anyValid) atomicOr(anyValid, 1u);
}

// One-workgroup sort of up to kSmallSortMax new boxes by (comparable key, index): a bitonic
// network in LDS. Equal comparable keys mean equal boxes, so this is the stable order by
// (key ascending, deeper first). Writes the boxes in that order.
constexpr uint32_t kSmallSortMax = 4096, kSmallSortThreads = 1024;
static __global__ __launch_bounds__(kSmallSortThreads) void smallBoxSortKernel(const uint64_t* k, const uint8_t* d,
                                                                              const uint8_t* f, uint32_t m,
                                                                              uint64_t* ok, uint8_t* od,
                                                                              uint8_t* of) {
  __shared__ uint64_t sk[kSmallSortMax];
  __shared__ uint32_t si[kSmallSortMax];
  uint32_t n2 = 1;
  while (n2 < m) n2 <<= 1;
  for (uint32_t i = threadIdx.x; i < n2; i += kSmallSortThreads) {
    sk[i] = i < m ? k[i] : ~0ull;
    si[i] = i;  // padding (index >= m) sorts last
  }
  __syncthreads();
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < n2 / 2; t += kSmallSortThreads) {
        const uint32_t a = 2 * t - (t & (stride - 1));  // first of the pair
        const uint32_t b = a + stride;
        const bool up = (a & size) == 0;
        const uint64_t ka = sk[a], kb = sk[b];
        const uint32_t ia = si[a], ib = si[b];
        const bool gt = ka > kb || (ka == kb && ia > ib);
        if (gt == up) {
          sk[a] = kb;
          sk[b] = ka;
          si[a] = ib;
          si[b] = ia;
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < m; i += kSmallSortThreads) {
    const uint32_t j = si[i];
    ok[i] = sk[i];
    od[i] = d[j];
    of[i] = f[j];
  }
}

// the new boxes, compacted: at leaf depth one level below the deepest mutual ancestor with the
// leaves on either side of the lower bound (AdaptDepthOfFirstBoxFunctor, Functors.hpp:379-389),
// otherwise truncated to d
static __global__ __launch_bounds__(kBlock) void insertNewKernel(const uint64_t* pk, const uint32_t* pos,
                                                                uint32_t total, uint64_t P, int leafForm, uint8_t d,
                                                                const uint64_t* key, const uint8_t* depth, uint32_t n,
                                                                uint8_t flags, uint64_t* outKey, uint8_t* outDepth,
                                                                uint8_t* outFlags) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= P) return;
  // pos is the exclusive scan of the compaction flags: point i has a slot iff the next offset
  // (the total after the last point) is larger
  const uint32_t next = i + 1 < P ? pos[i + 1] : total;
  if (next == pos[i]) return;
  const uint64_t p = pk[i];  // the point's box at depth 64: its comparable key is its pattern
  int nd;
  if (leafForm) {
    uint32_t lo = 0, len = n;
    while (len > 0) {
      const uint32_t half = len >> 1;
      if (key[lo + half] < p) {
        lo += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    const uint32_t i1 = lo > 0 ? lo - 1 : 0, i2 = lo < n - 1 ? lo : n - 1;
    const int m1 = mutualAncestorDepth(p, kMaxDepth, bitsAt(key[i1], depth[i1]), depth[i1]);
    const int m2 = mutualAncestorDepth(p, kMaxDepth, bitsAt(key[i2], depth[i2]), depth[i2]);
    nd = min(max(m1, m2) + 1, int(kMaxDepth));
  } else {
    nd = d;
  }
  const uint64_t b = nd >= kMaxDepth ? p : (p >> (kMaxDepth - nd));
  const uint32_t o = pos[i];
  outKey[o] = keyOf(b, nd);
  outDepth[o] = uint8_t(nd);
  outFlags[o] = flags;
}

// sort keys of the boxes (comparable key ascending, depth descending) by part: 0 = 255 - depth,
// 1 = low 32 bits of the key, 2 = high 32 bits, each gathered through idx
static __global__ __launch_bounds__(kBlock) void boxSortKeyKernel(const uint64_t* key, const uint8_t* depth,
                                                                 const uint32_t* idx, uint64_t n, int part,
                                                                 uint32_t* out) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = idx[i];
  out[i] = part == 0 ? uint32_t(255 - depth[j]) : (part == 1 ? uint32_t(key[j]) : uint32_t(key[j] >> 32));
}

static __global__ __launch_bounds__(kBlock) void gatherBoxesKernel(const uint64_t* key, const uint8_t* depth,
                                                                  const uint8_t* flags, const uint32_t* idx,
                                                                  uint64_t n, uint64_t* ok, uint8_t* od,
                                                                  uint8_t* of) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = idx[i];
  ok[i] = key[j];
  od[i] = depth[j];
  of[i] = flags[j];
}

// the order of the insertion's sort: comparable key ascending, deeper first on equal keys
__device__ inline bool boxLess(uint64_t ka, uint8_t da, uint64_t kb, uint8_t db) {
  return ka < kb || (ka == kb && da > db);
}

// Merge of the m sorted new boxes into the n sorted leaves, in the order a stable sort of the
// leaves followed by the new boxes gives: new box j goes after every leaf not greater than it
// (mergeNewKernel), leaf i after every new box less than it (mergeLeafKernel).
static __global__ __launch_bounds__(kBlock) void mergeNewKernel(const uint64_t* nk, const uint8_t* nd,
                                                               const uint8_t* nf, uint32_t m, const uint64_t* lk,
                                                               const uint8_t* ld, uint32_t n, uint64_t* ok,
                                                               uint8_t* od, uint8_t* of) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= m) return;
  const uint64_t k = nk[j];
  const uint8_t d = nd[j];
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (!boxLess(k, d, lk[lo + half], ld[lo + half])) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const uint64_t o = uint64_t(j) + lo;
  ok[o] = k;
  od[o] = d;
  of[o] = nf[j];
}

static __global__ __launch_bounds__(kBlock) void mergeLeafKernel(const uint64_t* lk, const uint8_t* ld,
                                                                const uint8_t* lf, uint32_t n, const uint64_t* nk,
                                                                const uint8_t* nd, uint32_t m, uint64_t* ok,
                                                                uint8_t* od, uint8_t* of) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = lk[i];
  const uint8_t d = ld[i];
  uint32_t lo = 0, len = m;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (boxLess(nk[lo + half], nd[lo + half], k, d)) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const uint64_t o = uint64_t(i) + lo;
  ok[o] = k;
  od[o] = d;
  of[o] = lf[i];
}

// runs of consecutive boxes on one path (hasSamePath)
static __global__ __launch_bounds__(kBlock) void pathRunKernel(const uint64_t* key, const uint8_t* depth, uint32_t n,
                                                              uint32_t* start, uint8_t* startFlag) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  bool st = i == 0;
  if (!st) st = !samePath(bitsAt(key[i - 1], depth[i - 1]), depth[i - 1], bitsAt(key[i], depth[i]), depth[i]);
  start[i] = st ? 1u : 0u;
  startFlag[i] = st ? 1 : 0;
}

// each run's box: the largest pattern, the largest depth, the OR of the flags
// (ComputeMaxBitPatternMaxDepthOrFlagsFunctor, Functors.hpp:391-401)
static __global__ __launch_bounds__(kBlock) void pathReduceKernel(const uint64_t* key, const uint8_t* depth,
                                                                 const uint8_t* flags, const uint32_t* run,
                                                                 uint32_t n, unsigned long long* rb, uint32_t* rd,
                                                                 uint32_t* rf) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = run[i];
  atomicMax(rb + r, (unsigned long long)bitsAt(key[i], depth[i]));
  atomicMax(rd + r, uint32_t(depth[i]));
  atomicOr(rf + r, uint32_t(flags[i]));
}

// a run's box is kept unless it is on one path with the previous run's box (unique, comparing
// neighbours)
static __global__ __launch_bounds__(kBlock) void pathKeepKernel(const unsigned long long* rb, const uint32_t* rd,
                                                               uint32_t m, uint32_t* keep) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  keep[t] = (t == 0 || !samePath(rb[t - 1], int(rd[t - 1]), rb[t], int(rd[t]))) ? 1u : 0u;
}

static __global__ __launch_bounds__(kBlock) void pathWriteKernel(const unsigned long long* rb, const uint32_t* rd,
                                                                const uint32_t* rf, const uint32_t* pos, uint32_t m,
                                                                TreeView out) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  const bool kp = t == 0 || !samePath(rb[t - 1], int(rd[t - 1]), rb[t], int(rd[t]));
  if (!kp) return;
  const uint32_t o = pos[t];
  out.key[o] = keyOf(rb[t], int(rd[t]));
  out.depth[o] = uint8_t(rd[t]);
  out.flags[o] = uint8_t(rf[t]);
}

// boxes(depth) rows {centre[DIM], radius[DIM], flags} (ImplicitBoxTree.hpp:611-775,
// coordinateFunctions.hpp:95-159): one per leaf (start == nullptr), or one per box at depth d
// from the first leaf of the box, truncated to d, with the flags combined over the box (acc)
template <int DIM>
static __global__ __launch_bounds__(kBlock) void boxesKernel(Domain dom, TreeView v, uint32_t n, uint8_t d,
                                                            const uint8_t* start, const uint32_t* box,
                                                            const uint32_t* acc, double* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint64_t key;
  int dep;
  uint8_t f;
  uint64_t row;
  if (start) {
    if (!start[i]) return;
    uint64_t b;
    truncateTo(v.key[i], v.depth[i], d, b, dep);
    key = keyOf(b, dep);
    f = uint8_t(acc[box[i]]);
    row = box[i];
  } else {
    key = v.key[i];
    dep = v.depth[i];
    f = v.flags[i];
    row = i;
  }
  double c[DIM], r[DIM];
  boxCoords<DIM>(dom, key, dep, c, r);
  double* o = out + row * (2 * DIM + 1);
#pragma unroll
  for (int j = 0; j < DIM; ++j) {
    o[j] = c[j];
    o[DIM + j] = r[j];
  }
  o[2 * DIM] = double(f);
}

static __global__ __launch_bounds__(kBlock) void depthSourceCountKernel(const uint8_t* startFlag, const uint32_t* box,
                                                                       const uint32_t* ok, uint32_t n, uint32_t* pick) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  pick[i] = (startFlag[i] && (!ok || ok[box[i]])) ? 1u : 0u;
}

static __global__ __launch_bounds__(kBlock) void depthSourceWriteKernel(TreeView v, uint32_t n, uint8_t d,
                                                                       const uint8_t* startFlag, const uint32_t* box,
                                                                       const uint32_t* ok, const uint32_t* pos,
                                                                       uint64_t* srcKey, uint8_t* srcDepth) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !(startFlag[i] && (!ok || ok[box[i]]))) return;
  uint64_t b;
  int td;
  truncateTo(v.key[i], v.depth[i], d, b, td);
  srcKey[pos[i]] = keyOf(b, td);
  srcDepth[pos[i]] = uint8_t(td);
}

// ---------------------------------------------------------------------------------------------
// Transition matrix rows (ImplicitBoxTree.hpp:2932-3159): the k targets of each source box are
// sorted inside a wave segment of S = next_pow2(k) lanes by a bitonic network over shuffles and
// run-length encoded; a row's entries come out in ascending column order (row-major COO).
// pass 0 writes per-row entry counts, pass 1 writes (column, count) at rowPtr[row].
// ---------------------------------------------------------------------------------------------
struct TmRowArgs {
  const uint32_t* col;   // P targets, box-major
  uint64_t nRows;
  uint32_t k;
  int keepInvalid;
  uint32_t invalidLabel;
  uint32_t* rowCount;    // pass 0
  const uint32_t* rowPtr;  // pass 1
  uint32_t* outCol;
  uint32_t* outCnt;
  uint32_t* anyInvalid;
  int pass;
};

template <int S>
static __global__ __launch_bounds__(kBlock) void tmRowsKernel(TmRowArgs a) {
  constexpr int SEGS = kWave / S;
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / S, pos = lane % S;
  const uint64_t waveId = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / kWave;
  const uint64_t nWaves = uint64_t(gridDim.x) * kBlock / kWave;
  for (uint64_t grp = waveId; grp * SEGS < a.nRows; grp += nWaves) {
    const uint64_t row = grp * SEGS + seg;
    const bool realRow = row < a.nRows;
    uint32_t v = 0xFFFFFFFFu;
    if (realRow && uint32_t(pos) < a.k) v = a.col[row * a.k + pos];
    // bitonic sort within the segment
#pragma unroll
    for (int size = 2; size <= S; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const uint32_t o = __shfl_xor(v, stride, kWave);
        const bool up = (pos & size) == 0;
        const bool lower = (pos & stride) == 0;
        v = (lower == up) ? min(v, o) : max(v, o);
      }
    }
    const uint32_t prev = __shfl_up(v, 1, kWave);
    const bool start = (pos == 0) || (v != prev);
    const uint64_t startMask = __ballot(start);
    const uint64_t segMask = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << (seg * S));
    // next run start in this segment (or segment end)
    const uint64_t aboveMask = (lane == 63) ? 0ull : (startMask & segMask & ~((1ull << (lane + 1)) - 1ull));
    const int next = aboveMask ? __builtin_ctzll(aboveMask) : (seg * S + S);
    uint32_t cnt = uint32_t(next - lane);
    bool entry = start && realRow;
    bool invalidHere = false;
    if (v == 0xFFFFFFFFu) {
      cnt -= uint32_t(S) - a.k;  // padding lanes sort last, after the invalid targets
      invalidHere = entry && cnt > 0;
      entry = invalidHere && a.keepInvalid;
    }
    // one atomic per wave that saw an invalid target (same-address atomics serialise chip-wide)
    if (a.pass == 0 && __ballot(invalidHere) && lane == 0) atomicOr(a.anyInvalid, 1u);
    const uint64_t entryMask = __ballot(entry) & segMask;
    if (a.pass == 0) {
      if (pos == 0 && realRow) a.rowCount[row] = uint32_t(__popcll(entryMask));
    } else if (entry) {
      const uint64_t below = entryMask & ((1ull << lane) - 1ull);
      const uint32_t dst = a.rowPtr[row] + uint32_t(__popcll(below));
      a.outCol[dst] = (v == 0xFFFFFFFFu) ? a.invalidLabel : v;
      a.outCnt[dst] = cnt;
    }
  }
}

// General transition-matrix rows for any k: (row, col) pairs sorted row-major by two stable radix
// passes, then run-length encoded here. count(i) = 1 at the start of a run that becomes an entry.
struct RunLengthOp : OpBase {
  const uint32_t* row;
  const uint32_t* col;
  uint64_t n;
  int keepInvalid;
  uint32_t invalidLabel;
  uint32_t* outRow;
  uint32_t* outCol;
  uint32_t* outCnt;
  __device__ bool isStart(uint32_t i) const { return i == 0 || row[i] != row[i - 1] || col[i] != col[i - 1]; }
  __device__ uint32_t count(uint32_t i) const {
    return (isStart(i) && (col[i] != 0xFFFFFFFFu || keepInvalid)) ? 1u : 0u;
  }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const {
    uint64_t j = i + 1;
    while (j < n && row[j] == row[i] && col[j] == col[i]) ++j;
    outRow[pos] = row[i];
    outCol[pos] = col[i] == 0xFFFFFFFFu ? invalidLabel : col[i];
    outCnt[pos] = uint32_t(j - i);
  }
};

static __global__ __launch_bounds__(kBlock) void rowOfPointKernel(uint32_t* row, uint64_t n, uint32_t k) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) row[i] = uint32_t(i / k);
}

static __global__ __launch_bounds__(kBlock) void anyInvalidKernel(const uint32_t* col, uint64_t n, uint32_t* flag) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  const bool inv = i < n && col[i] == 0xFFFFFFFFu;
  if (__ballot(inv) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, 1u);
}

// Expand CSR row pointers into the COO row index per entry.
static __global__ __launch_bounds__(kBlock) void expandRowsKernel(const uint32_t* rowPtr, uint64_t nRows, uint32_t* rows) {
  const uint64_t r = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= nRows) return;
  for (uint32_t e = rowPtr[r]; e < rowPtr[r + 1]; ++e) rows[e] = uint32_t(r);
}

// values = count / k (or count) as double (ImplicitBoxTree.hpp:3145-3158)
static __global__ __launch_bounds__(kBlock) void valuesKernel(const uint32_t* cnt, uint64_t nnz, double k, int useAbs,
                                                       double* val) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nnz) return;
  const double c = double(cnt[e]);
  val[e] = useAbs ? c : c / k;
}

// ---------------------------------------------------------------------------------------------
// Stable LSD radix sort of uint32 keys with two uint32 payloads (8-bit digits). Block-local
// stable ranks come from wave ballots over the digit bits plus per-wave digit counts in LDS.
// hist layout: [digit * nBlocks + block], so one exclusive scan yields every output offset.
// ---------------------------------------------------------------------------------------------
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kBlock * kRadixItems;

static __global__ __launch_bounds__(kBlock) void radixHistKernel(const uint32_t* keys, uint64_t n, int shift,
                                                          uint32_t* hist, uint32_t nBlocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = uint64_t(blockIdx.x) * kRadixTile;
  for (int t = 0; t < kRadixItems; ++t) {
    const uint64_t i = base + uint64_t(t) * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[uint64_t(threadIdx.x) * nBlocks + blockIdx.x] = h[threadIdx.x];
}

static __global__ __launch_bounds__(kBlock) void radixScatterKernel(const uint32_t* keysIn, const uint32_t* p0In,
                                                             const uint32_t* p1In, uint64_t n, int shift,
                                                             const uint32_t* offsets, uint32_t nBlocks,
                                                             uint32_t* keysOut, uint32_t* p0Out, uint32_t* p1Out) {
  __shared__ uint32_t running[256];
  __shared__ uint32_t waveCnt[kBlock / kWave][256];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  running[threadIdx.x] = offsets[uint64_t(threadIdx.x) * nBlocks + blockIdx.x];
  const uint64_t base = uint64_t(blockIdx.x) * kRadixTile;
  for (int t = 0; t < kRadixItems; ++t) {
    for (int w = 0; w < kBlock / kWave; ++w) waveCnt[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = base + uint64_t(t) * kBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0, p0 = 0, p1 = 0, dg = 0;
    if (valid) {
      key = keysIn[i];
      p0 = p0In[i];
      p1 = p1In[i];
      dg = (key >> shift) & 0xFFu;
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t m = __ballot((dg >> bit) & 1u);
      peers &= ((dg >> bit) & 1u) ? m : ~m;
    }
    const uint32_t rankInWave = uint32_t(__popcll(peers & ((1ull << lane) - 1ull)));
    const int leader = peers ? __builtin_ctzll(peers) : -1;
    if (valid && lane == leader) waveCnt[wave][dg] = uint32_t(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int w = 0; w < wave; ++w) before += waveCnt[w][dg];
      const uint32_t dst = running[dg] + before + rankInWave;
      keysOut[dst] = key;
      p0Out[dst] = p0;
      p1Out[dst] = p1;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kBlock / kWave; ++w) add += waveCnt[w][threadIdx.x];
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

// Exclusive scan of a large uint32 array in place (three phases over kTile-element blocks).
struct ScanArrayOp : OpBase {
  uint32_t* v;
  __device__ uint32_t count(uint32_t i) const { return v[i]; }
  __device__ void emit(uint32_t, uint32_t, uint32_t) const {}
};

static __global__ __launch_bounds__(kBlock) void scanArrayScatterKernel(uint32_t* v, uint32_t n, const uint32_t* blockOffsets) {
  const uint64_t base = uint64_t(blockIdx.x) * kTile + uint64_t(threadIdx.x) * kItems;
  uint32_t c[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int t = 0; t < kItems; ++t) {
    const uint64_t i = base + t;
    c[t] = i < n ? v[i] : 0u;
    s += c[t];
  }
  uint32_t tot;
  uint32_t pos = blockExclusiveScan(s, tot) + blockOffsets[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kItems; ++t) {
    const uint64_t i = base + t;
    if (i < n) v[i] = pos;
    pos += c[t];
  }
}

// offsets[v] = number of entries with idx < v, for sorted idx (compressIndexVector,
// CooMatrix.hpp:14-56), v = 0..dim
static __global__ __launch_bounds__(kBlock) void compressKernel(const uint32_t* idx, uint64_t nnz, uint32_t dim,
                                                         uint32_t* out) {
  const uint64_t v = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v > dim) return;
  uint64_t lo = 0, len = nnz;
  while (len > 0) {
    const uint64_t half = len >> 1;
    if (idx[lo + half] < v) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  out[v] = uint32_t(lo);
}

// ---------------------------------------------------------------------------------------------
// Power iteration on the CSC form (column j holds the entries P_ij in ascending row order):
// y_j = sum_i mu_i P_ij summed in ascending row order, then mu <- y / ||y||_1. Block partial
// sums go to fixed slots and are reduced in a fixed order, so the result is deterministic.
// ---------------------------------------------------------------------------------------------
__device__ inline double blockReduceSumF64(double v) {
  __shared__ double waveSums[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  if (lane == 0) waveSums[wave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) t += waveSums[w];
  __syncthreads();
  return t;
}

// One launch per iteration t = 1, 2, ...: y_t = mu_{t-1} P, where mu_{t-1} = y_{t-1} / ||y_{t-1}||_1
// is recomputed on the fly from the previous iterate (the same IEEE quotient the oracle stores),
// together with the residual ||mu_{t-1} - mu_{t-2}||_1 of the previous iteration; block partials
// of both sums go to fixed slots for piReduceKernel. y rotates over 3 slots.
struct PiState {
  double norm[3];
  double resid;
  uint32_t t;
  uint32_t pad;
  int32_t done;
  uint32_t final;
};

// kPiLanes lanes share a column: they load and multiply its entries in parallel, and the first
// lane of the group adds the products in ascending-row order (the oracle's order) via shuffles.
constexpr int kPiLanes = 8;

static __global__ __launch_bounds__(kBlock) void piStepKernel(const uint32_t* __restrict__ colPtr,
                                                       const uint32_t* __restrict__ rowIdx,
                                                       const double* __restrict__ val, const double* __restrict__ mu0,
                                                       double* y, uint32_t n, double* partAbs, double* partRes,
                                                       const PiState* st) {
  if (st->done) return;
  const uint32_t t = st->t;
  const uint32_t s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;  // slots of t, t-1, t-2
  const double* yPrev = y + uint64_t(s1) * n;
  const double* yPrev2 = y + uint64_t(s2) * n;
  double* yCur = y + uint64_t(s0) * n;
  const double nPrev = st->norm[s1], nPrev2 = st->norm[s2];
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane % kPiLanes, lead = lane - sub;
  const uint32_t j = (blockIdx.x * kBlock + threadIdx.x) / kPiLanes;
  uint32_t e0 = 0, e1 = 0;
  if (j < n) {
    e0 = colPtr[j];
    e1 = colPtr[j + 1];
  }
  // trip count uniform across the wave so that every lane takes part in every shuffle
  uint32_t trips = (e1 - e0 + kPiLanes - 1) / kPiLanes;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) trips = max(trips, (uint32_t)__shfl_xor(trips, o, kWave));
  double s = 0.0;
  for (uint32_t r = 0; r < trips; ++r) {
    const uint32_t e = e0 + r * kPiLanes + sub;
    double prod = 0.0;
    if (e < e1) {
      const uint32_t i = rowIdx[e];
      const double m = (t == 1) ? mu0[i] : yPrev[i] / nPrev;
      prod = m * val[e];
    }
#pragma unroll
    for (int k = 0; k < kPiLanes; ++k) {
      const double pk = __shfl(prod, lead + k, kWave);
      if (e0 + r * kPiLanes + k < e1) s += pk;
    }
  }
  double sAbs = 0.0, sRes = 0.0;
  if (j < n && sub == 0) {
    yCur[j] = s;
    sAbs = fabs(s);
    if (t >= 2) {
      const double a = yPrev[j] / nPrev;
      const double b = (t == 2) ? mu0[j] : yPrev2[j] / nPrev2;
      sRes = fabs(a - b);
    }
  }
  sAbs = blockReduceSumF64(sAbs);
  sRes = blockReduceSumF64(sRes);
  if (threadIdx.x == 0) {
    partAbs[blockIdx.x] = sAbs;
    partRes[blockIdx.x] = sRes;
  }
}

// Sums the block partials of piStepKernel in block order (deterministic) and decides whether
// iteration t-1 ends the loop. One block; runs after every piStepKernel.
static __global__ __launch_bounds__(kBlock) void piReduceKernel(const double* partAbs, const double* partRes, uint32_t nb,
                                                         PiState* st, uint32_t maxIter, double tol) {
  if (st->done) return;
  const uint32_t t = st->t;
  // issue all of a thread's loads before summing them (one memory latency instead of one per
  // loop trip); the summation order per thread stays fixed
  constexpr int U = 8;
  double ta = 0.0, tr = 0.0;
  for (uint32_t b0 = 0; b0 < nb; b0 += kBlock * U) {
    double va[U], vr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t b = b0 + u * kBlock + threadIdx.x;
      va[u] = b < nb ? partAbs[b] : 0.0;
      vr[u] = b < nb ? partRes[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ta += va[u];
      tr += vr[u];
    }
  }
  ta = blockReduceSumF64(ta);
  tr = blockReduceSumF64(tr);
  if (threadIdx.x == 0) {
    st->norm[t % 3] = ta;
    const uint32_t completed = t - 1;
    if (completed >= 1) {
      st->resid = tr;
      if ((tol >= 0.0 && tr <= tol) || completed >= maxIter) {
        st->done = 1;
        st->final = completed;
      }
    }
    st->t = t + 1;
  }
}

// Single-kernel form for grids of at most kPiFusedMaxBlocks blocks of kPiBlock threads: launch t
// first re-derives norm_{t-1} and the residual of iteration t-2 from the previous launch's block
// partials (every block sums them in the same order, so all blocks agree), decides whether
// iteration t-2 ended the loop, then computes y_t and its partials into the other parity slot.
// Each block counts its own launches (blockIter[b]), so no block waits on or races with another.
constexpr int kPiBlock = 1024;
constexpr uint32_t kPiFusedMaxBlocks = 512;

__device__ inline double blockSumF64Big(double v) {
  __shared__ double ws[kPiBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  if (lane == 0) ws[wave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kPiBlock / kWave; ++w) t += ws[w];
  __syncthreads();
  return t;
}

static __global__ __launch_bounds__(kPiBlock) void piFusedKernel(const uint32_t* __restrict__ colPtr,
                                                          const uint32_t* __restrict__ rowIdx,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ mu0, double* y, uint32_t n,
                                                          double* part /* [2][2][nb] */, uint32_t* blockIter,
                                                          PiState* st, uint32_t maxIter, double tol) {
  const uint32_t nb = gridDim.x;
  const uint32_t t = blockIter[blockIdx.x] + 1;  // this block's launch index
  if (st->done) return;
  double nPrev = 0.0, rPrev = -1.0;
  if (t >= 2) {
    const double* pa = part + uint64_t((t - 1) & 1) * 2 * nb;
    double a = 0.0, r = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += kPiBlock) {
      a += pa[b];
      r += pa[nb + b];
    }
    nPrev = blockSumF64Big(a);
    rPrev = blockSumF64Big(r);
  }
  const uint32_t completed = t >= 2 ? t - 2 : 0;  // iteration whose residual is rPrev
  if (completed >= 1 && ((tol >= 0.0 && rPrev <= tol) || completed >= maxIter)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->final = completed;
      st->resid = rPrev;
    }
    return;
  }
  if (threadIdx.x == 0) blockIter[blockIdx.x] = t;
  if (blockIdx.x == 0 && threadIdx.x == 0 && t >= 2) {
    st->norm[(t - 1) % 3] = nPrev;
    if (completed >= 1) st->resid = rPrev;
  }
  const uint32_t s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;
  const double* yPrev = y + uint64_t(s1) * n;
  const double* yPrev2 = y + uint64_t(s2) * n;
  double* yCur = y + uint64_t(s0) * n;
  const double nPrev2 = t >= 3 ? st->norm[s2] : 0.0;
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane % kPiLanes, lead = lane - sub;
  double sAbs = 0.0, sRes = 0.0;
  const uint64_t total = uint64_t(n) * kPiLanes;
  const uint64_t stride = uint64_t(nb) * kPiBlock;
  for (uint64_t g = uint64_t(blockIdx.x) * kPiBlock + threadIdx.x; g - lane < total; g += stride) {
    const uint32_t j = uint32_t(g / kPiLanes);
    uint32_t e0 = 0, e1 = 0;
    if (g < total) {
      e0 = colPtr[j];
      e1 = colPtr[j + 1];
    }
    uint32_t trips = (e1 - e0 + kPiLanes - 1) / kPiLanes;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) trips = max(trips, (uint32_t)__shfl_xor(trips, o, kWave));
    double s = 0.0;
    for (uint32_t r = 0; r < trips; ++r) {
      const uint32_t e = e0 + r * kPiLanes + sub;
      double prod = 0.0;
      if (e < e1) {
        const uint32_t i = rowIdx[e];
        const double m = (t == 1) ? mu0[i] : yPrev[i] / nPrev;
        prod = m * val[e];
      }
#pragma unroll
      for (int k = 0; k < kPiLanes; ++k) {
        const double pk = __shfl(prod, lead + k, kWave);
        if (e0 + r * kPiLanes + k < e1) s += pk;
      }
    }
    if (g < total && sub == 0) {
      yCur[j] = s;
      sAbs += fabs(s);
      if (t >= 2) {
        const double a = yPrev[j] / nPrev;
        const double b = (t == 2) ? mu0[j] : yPrev2[j] / nPrev2;
        sRes += fabs(a - b);
      }
    }
  }
  sAbs = blockSumF64Big(sAbs);
  sRes = blockSumF64Big(sRes);
  if (threadIdx.x == 0) {
    double* pc = part + uint64_t(t & 1) * 2 * nb;
    pc[blockIdx.x] = sAbs;
    pc[nb + blockIdx.x] = sRes;
  }
}

// mu = y_final / ||y_final||_1 (or mu0 when no iteration ran)
// Persistent form of the loop: one launch whose grid fits the device at once (every block is
// resident for the whole run on an otherwise idle device).
// Per iteration each block computes y_t for its columns (one lane per column, products added in
// ascending-row order as the oracle does) and publishes its partial sums of |y_t| and of the
// previous residual in its slot of buffer t % 3. A slot holds kCoopEmpty until written. Block 0
// polls every slot of the buffer until none is empty, sums them in block order and publishes the
// totals in result slot t % 3 (kPiResCopies copies); the other blocks poll their copy. So all
// blocks derive the same norm and the same stopping decision. A block clears its partial slot of
// buffer (t+1) % 3 at the start of iteration t, and block 0 the result slot (t+1) % 3: both were
// last read in iteration t-2, which every block has finished once the wait of iteration t-1 is
// over.
// y rotates over 3 slots (y_t in slot t % 3) and a lane clears its columns of slot (t+1) % 3
// (y_{t-2}, read for the last time in iteration t-1) to kCoopEmpty in iteration t, before its
// block publishes. So once a block has seen the totals of iteration t-1, slot t % 3 holds
// kCoopEmpty or y_t, never an older iterate. That lets each lane gather the y_t entries of its
// first column right after its block publishes y_t, while the totals are still on their way:
// the gathers overlap the wait, and a lane re-reads the entries that came back empty (their
// owners had not stored them yet) until none is left. Everything else is gathered after the
// wait, when every y_t entry is in place.
// Every wait is bounded: past spinLimit polls a block raises *err and leaves, and the others
// leave at their next wait (the host then recomputes with the batched form).
// Coherence across XCDs: everything written during the run (y, partials, totals) is written and
// read with agent-scope relaxed atomics (coherent without L2 write-back or invalidation), and
// each wave waits for its stores to complete before its block publishes. The matrix is read-only
// and stays cached in L2 across iterations.
constexpr uint32_t kCoopMaxBlocks = 512;
constexpr int kCoopUnroll = 8;
constexpr int kCoopCache = 3 * kCoopUnroll;  // entries of a lane's first column held in registers
constexpr uint64_t kCoopEmpty = 0xFFF4B12B12B12B12ull;  // a NaN pattern arithmetic never produces
// block 0 writes its totals to kPiResCopies copies, one 128-byte line each, and block b polls
// copy b % kPiResCopies, so that no single line takes every block's polls
constexpr uint32_t kPiResCopies = 16, kPiResStride = 16;
constexpr uint32_t kPiResWords = 3 * kPiResCopies * kPiResStride;

struct PiCoopArgs {
  const uint32_t* colPtr;
  const uint32_t* rowIdx;
  const double* val;
  const double* mu0;
  double* y;          // 3 * n, all kCoopEmpty at launch
  uint64_t* part;     // 3 buffers * gridDim.x slots * 2, all kCoopEmpty at launch
  uint64_t* res;      // 3 slots * kPiResCopies copies * kPiResStride (block 0's totals), all kCoopEmpty
  PiState* st;
  int* err;           // zero at launch
  uint32_t n;
  uint32_t maxIter;
  double tol;
  uint32_t spinLimit;
};

template <class T>
__device__ inline T ldAgent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void stAgent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// y entries as bit patterns (the empty pattern is compared bitwise)
__device__ inline uint64_t ldBits(const double* p) { return ldAgent(reinterpret_cast<const uint64_t*>(p)); }
__device__ inline void stBits(double* p, uint64_t v) { stAgent(reinterpret_cast<uint64_t*>(p), v); }

// y entry at a byte offset below 2^32 from a uniform base (one 32-bit offset register per gather)
__device__ inline uint64_t ldAt(const double* base, uint32_t off) {
  return ldAgent(reinterpret_cast<const uint64_t*>(reinterpret_cast<const char*>(base) + off));
}

// the cached entries of a lane's first column (byte offsets ci), gathered from src: groups of
// kCoopUnroll gathers, each issued where some column of the wave reaches it (clamped indices
// keep the others in bounds), all in flight together
__device__ inline void piGatherCached(const double* src, const uint32_t (&ci)[kCoopCache], uint32_t c0, uint32_t c1,
                                      uint64_t (&m)[kCoopCache]) {
#pragma unroll
  for (int g = 0; g < kCoopCache; g += kCoopUnroll) {
    if (__any(c0 + g < c1)) {  // wave-uniform: no wait at the join
#pragma unroll
      for (int u = g; u < g + kCoopUnroll; ++u) m[u] = ldAt(src, ci[u]);
    }
  }
}

// true where a cached entry the column uses came back empty
__device__ inline bool piMissing(const uint64_t (&m)[kCoopCache], uint32_t c0, uint32_t c1) {
  bool miss = false;
#pragma unroll
  for (int u = 0; u < kCoopCache; ++u) miss = miss || (c0 + u < c1 && m[u] == kCoopEmpty);
  return miss;
}

// re-read the empty entries (all of a lane's re-reads in flight together)
__device__ inline void piRegather(const double* src, const uint32_t (&ci)[kCoopCache], uint32_t c0, uint32_t c1,
                                  uint64_t (&m)[kCoopCache]) {
#pragma unroll
  for (int u = 0; u < kCoopCache; ++u)
    if (c0 + u < c1 && m[u] == kCoopEmpty) m[u] = ldAt(src, ci[u]);
}

// SLOTS: partial slots per polling lane of block 0 (grids up to SLOTS * 64 blocks)
template <uint32_t SLOTS>
static __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void piCoopKernel(PiCoopArgs a) {
  constexpr uint32_t kSlots = SLOTS;
  constexpr int kWaves = kBlock / kWave;
  const uint32_t G = gridDim.x, b = blockIdx.x, tid = threadIdx.x, n = a.n;
  const uint32_t lane = tid & (kWave - 1), wv = tid / kWave;
  __shared__ double wsum[2][kWaves];
  __shared__ double sums[2];
  __shared__ int leave, failed;
  if (a.st->done) return;
  uint32_t t = a.st->t;
  double nPrev = t >= 2 ? a.st->norm[(t + 2) % 3] : 1.0;  // norm of y_{t-1}
  double nPrev2 = t >= 3 ? a.st->norm[(t + 1) % 3] : 1.0;  // norm of y_{t-2}
  // the lane's first column keeps its first kCoopCache entries (row byte offsets and values) in
  // registers for the whole run (indices clamped into the column; an empty column holds row 0
  // and adds nothing)
  const uint32_t p0 = b * kBlock + tid;
  uint32_t c0 = 0, c1 = 0;
  uint32_t ci[kCoopCache];
  double cv[kCoopCache];
#pragma unroll
  for (int u = 0; u < kCoopCache; ++u) {
    ci[u] = 0;
    cv[u] = 0.0;
  }
  // the first column's own y_{t-1} and y_{t-2} (mu_0 stands for y_0, with norm 1)
  double own1 = 0.0, own2 = 0.0;
  if (p0 < n) {
    c0 = a.colPtr[p0];
    c1 = a.colPtr[p0 + 1];
    if (c1 > c0) {
#pragma unroll
      for (int u = 0; u < kCoopCache; ++u) {
        const uint32_t eu = min(c0 + u, c1 - 1);
        ci[u] = a.rowIdx[eu] * uint32_t(sizeof(double));  // n < 2^29
        cv[u] = a.val[eu];
      }
    }
    own1 = t >= 2 ? __longlong_as_double((long long)ldBits(a.y + uint64_t((t + 2) % 3) * n + p0)) : a.mu0[p0];
    own2 = t >= 3 ? __longlong_as_double((long long)ldBits(a.y + uint64_t((t + 1) % 3) * n + p0))
                  : (t == 2 ? a.mu0[p0] : 0.0);
  }
  // the first gathers: y_{t-1} (mu_0 in iteration 1) is complete at launch
  uint64_t m[kCoopCache];
#pragma unroll
  for (int u = 0; u < kCoopCache; ++u) m[u] = 0;
  piGatherCached((t == 1) ? a.mu0 : a.y + uint64_t((t + 2) % 3) * n, ci, c0, c1, m);
  for (;;) {
    if (tid == 0) {
      uint64_t* nextSlot = a.part + (uint64_t((t + 1) % 3) * G + b) * 2;
      stAgent(nextSlot, kCoopEmpty);
      stAgent(nextSlot + 1, kCoopEmpty);
      failed = 0;
    }
    if (b == 0 && tid < kPiResCopies) {
      uint64_t* rn = a.res + (((t + 1) % 3) * kPiResCopies + tid) * kPiResStride;
      stAgent(rn, kCoopEmpty);
      stAgent(rn + 1, kCoopEmpty);
    }
    const uint32_t s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;
    const double* yPrev = a.y + uint64_t(s1) * n;
    double* yOld = a.y + uint64_t(s2) * n;  // y_{t-2}: cleared below
    double* yCur = a.y + uint64_t(s0) * n;
    // iteration 1 multiplies by mu_0 itself: mu_0[i] / 1.0 is mu_0[i] exactly
    const double* src = (t == 1) ? a.mu0 : yPrev;
    const double div = (t == 1) ? 1.0 : nPrev;
    const double div2 = (t == 2) ? 1.0 : nPrev2;
    double sAbs = 0.0, sRes = 0.0;
    for (uint32_t p = p0; p < n; p += G * kBlock) {
      double s = 0.0, yp, yp2;
      uint32_t e, e1;
      if (p == p0) {
        // mu_{t-1}[i] = y_{t-1}[i] / ||y_{t-1}||_1 as the oracle rounds it, then the product
#pragma unroll
        for (int u = 0; u < kCoopCache; ++u) {
          const double prod = (__longlong_as_double((long long)m[u]) / div) * cv[u];
          s = (c0 + u < c1) ? s + prod : s;
        }
        e = c0 + kCoopCache;
        e1 = c1;
        yp = own1;
        yp2 = own2;
      } else {
        yp = __longlong_as_double((long long)ldBits(yPrev + p));
        yp2 = __longlong_as_double((long long)ldBits(((t == 2) ? a.mu0 : yOld) + p));
        e = a.colPtr[p];
        e1 = a.colPtr[p + 1];
      }
      // branch-free body (entry indices clamped to the column, sums selected) so that all of a
      // lane's gathers are in flight at once; y_{t-1} is complete here
      for (; e < e1; e += kCoopUnroll) {
        uint32_t i[kCoopUnroll];
        double v[kCoopUnroll], mm[kCoopUnroll];
#pragma unroll
        for (int u = 0; u < kCoopUnroll; ++u) {
          const uint32_t eu = min(e + u, e1 - 1);
          i[u] = a.rowIdx[eu];
          v[u] = a.val[eu];
        }
#pragma unroll
        for (int u = 0; u < kCoopUnroll; ++u) mm[u] = ldAgent(src + i[u]);
#pragma unroll
        for (int u = 0; u < kCoopUnroll; ++u) {
          const double prod = (mm[u] / div) * v[u];
          s = (e + u < e1) ? s + prod : s;
        }
      }
      stBits(yCur + p, uint64_t(__double_as_longlong(s)));
      sAbs += fabs(s);
      if (t >= 2) sRes += fabs(yp / nPrev - yp2 / div2);
      if (p == p0) {
        own2 = own1;
        own1 = s;
      }
      stBits(yOld + p, kCoopEmpty);  // after yp2 was read; the next value stored here is y_{t+1}
    }
    // block sums: each wave by a fixed shuffle tree, then the waves in order
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      sAbs += __shfl_xor(sAbs, o, kWave);
      sRes += __shfl_xor(sRes, o, kWave);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's y stores and clears have completed
    if (lane == 0) {
      wsum[0][wv] = sAbs;
      wsum[1][wv] = sRes;
    }
    __syncthreads();
    uint64_t* buf = a.part + uint64_t(t % 3) * G * 2;
    if (tid == 0) {
      double A = 0.0, R = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        A += wsum[0][w];
        R += wsum[1][w];
      }
      stAgent(buf + 2 * b, uint64_t(__double_as_longlong(A)));
      stAgent(buf + 2 * b + 1, uint64_t(__double_as_longlong(R)));
    }
    // the next iteration's first-column gathers of y_t, overlapping the wait
    piGatherCached(yCur, ci, c0, c1, m);
    if (tid < kWave) {
      // Block 0 polls every partial slot, sums them in block order and publishes the totals in
      // result slot t % 3; the other blocks poll that one slot. A poll round issues all of its
      // loads (slots, the error flag and the wave's empty gathers) before testing any, so it
      // costs one memory round trip.
      uint64_t* rs = a.res + ((t % 3) * kPiResCopies + b % kPiResCopies) * kPiResStride;
      double ta = 0.0, tr = 0.0;
      bool ok = false;
      if (b == 0) {
        // lane l holds slots l, l + 64, ... ; sums in that order, then a fixed shuffle tree.
        // Groups past the grid are skipped by a uniform test; lanes past the last slot re-read
        // slot G - 1.
        uint64_t va[kSlots], vr[kSlots];
#pragma unroll
        for (uint32_t k = 0; k < kSlots; ++k) va[k] = vr[k] = 0;
        uint32_t lt = tid;  // opaque: slot offsets are recomputed here, not held across iterations
        asm volatile("" : "+v"(lt));
        for (uint32_t spin = 0; spin < a.spinLimit; ++spin) {
          const int e = ldAgent(a.err);
#pragma unroll
          for (uint32_t k = 0; k < kSlots; ++k) {
            if (k * kWave < G) {
              const uint32_t off = min(lt + k * kWave, G - 1) * 16u;  // byte offset of the slot
              const char* base = reinterpret_cast<const char*>(buf);
              va[k] = ldAgent(reinterpret_cast<const uint64_t*>(base + off));
              vr[k] = ldAgent(reinterpret_cast<const uint64_t*>(base + off + 8));
            }
          }
          bool mine = true;
#pragma unroll
          for (uint32_t k = 0; k < kSlots; ++k)
            if (k * kWave < G) mine = mine && va[k] != kCoopEmpty && vr[k] != kCoopEmpty;
          ok = __all(mine);
          if (ok || e) break;
          __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (uint32_t k = 0; k < kSlots; ++k)
          if (tid + k * kWave < G) {
            ta += __longlong_as_double((long long)va[k]);
            tr += __longlong_as_double((long long)vr[k]);
          }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
          ta += __shfl_xor(ta, o, kWave);
          tr += __shfl_xor(tr, o, kWave);
        }
        // the clears of result slot (t+1) % 3 (start of this iteration) have completed: the
        // wave waited for its stores before the block published
        if (tid < kPiResCopies && ok) {
          uint64_t* rc = a.res + ((t % 3) * kPiResCopies + tid) * kPiResStride;
          stAgent(rc, uint64_t(__double_as_longlong(ta)));
          stAgent(rc + 1, uint64_t(__double_as_longlong(tr)));
        }
      }
      {
        uint64_t ra = 0, rr = 0;
        for (uint32_t spin = 0; spin < a.spinLimit; ++spin) {
          const int e = ldAgent(a.err);
          if (b != 0) {
            ra = ldAgent(rs);
            rr = ldAgent(rs + 1);
          }
          if (spin > 0 && __any(piMissing(m, c0, c1))) piRegather(yCur, ci, c0, c1, m);
          const bool have = b == 0 ? ok : (ra != kCoopEmpty && rr != kCoopEmpty);
          const bool done = have && !__any(piMissing(m, c0, c1));
          if (done || e) {
            ok = done;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (b != 0) {
          ta = __longlong_as_double((long long)ra);
          tr = __longlong_as_double((long long)rr);
        }
      }
      if (tid == 0) {
        if (!ok) stAgent(a.err, 1);
        leave = ok ? 0 : 1;
        sums[0] = ta;
        sums[1] = tr;
      }
    } else {
      // the other waves complete their gathers of y_t (their owners store them before they
      // publish, so the entries fill in while the totals arrive)
      bool ok = false;
      for (uint32_t spin = 0; spin < a.spinLimit; ++spin) {
        if (!__any(piMissing(m, c0, c1))) {
          ok = true;
          break;
        }
        const int e = ldAgent(a.err);
        piRegather(yCur, ci, c0, c1, m);
        if (e) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (!ok && lane == 0) {
        stAgent(a.err, 1);
        failed = 1;
      }
    }
    __syncthreads();
    if (leave || failed) return;
    const double ta = sums[0], tr = sums[1];
    const uint32_t completed = t - 1;
    const bool done = completed >= 1 && ((a.tol >= 0.0 && tr <= a.tol) || completed >= a.maxIter);
    if (done) {
      if (b == 0 && tid == 0) {
        a.st->norm[t % 3] = ta;
        a.st->norm[(t + 2) % 3] = nPrev;
        a.st->norm[(t + 1) % 3] = nPrev2;
        a.st->resid = tr;
        a.st->final = completed;
        a.st->t = t + 1;
        a.st->done = 1;
      }
      return;
    }
    nPrev2 = nPrev;
    nPrev = ta;
    ++t;
    __syncthreads();  // sums[] is rewritten by the next wait
  }
}

// sel[i] = 1 where leaf i meets the Hénon-Heiles energy band (4-dimensional trees)
static __global__ __launch_bounds__(kBlock) void energyBandKernel(Domain dom, const uint64_t* key, const uint8_t* depth,
                                                                 uint32_t n, double emin, double emax, uint8_t* sel) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double c[4], r[4];
  boxCoords<4>(dom, key[i], depth[i], c, r);
  sel[i] = henonHeilesBand(c, r, emin, emax) ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// Subspace iteration for the m <= kSubspaceMax leading eigenvalues (build-defined; the reference
// leaves eigenvalues to MATLAB eigs, subdivisionByNim.m:30-59). The basis is stored box by box
// (X[i * m + c]), so one matrix entry gathers m contiguous values. Y = X P with
// y_j[c] = sum_i x_i[c] P_ij in ascending row order; the basis is re-orthonormalised by Cholesky
// QR applied twice. Dot products: each thread sums its indices i = t, t + T, ... (T = all
// threads of the grid) in order, the block sums its threads by a fixed tree, and one block sums
// the block partials in a fixed order.
// ---------------------------------------------------------------------------------------------
constexpr int kSubspaceMax = 8;
constexpr uint32_t kDotBlocks = 1024;

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// X_0: column 0 = 1/n, column c > 0 = (splitmix64(8 i + c) >> 11) 2^-53 * 2 - 1
static __global__ __launch_bounds__(kBlock) void subspaceStartKernel(double* X, uint32_t n, int m) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  X[uint64_t(i) * m] = 1.0 / double(n);
  for (int c = 1; c < m; ++c)
    X[uint64_t(i) * m + c] = double(splitmix64(8 * uint64_t(i) + uint64_t(c)) >> 11) * 0x1p-53 * 2.0 - 1.0;
}

// one lane per column j; kSpmmUnroll entries' gathers in flight at a time (indices clamped into
// the column, sums selected), products added in ascending row order
constexpr int kSpmmUnroll = 8;

template <int M>
static __global__ __launch_bounds__(kBlock) void spmmKernel(const uint32_t* __restrict__ colPtr,
                                                           const uint32_t* __restrict__ rows,
                                                           const double* __restrict__ vals,
                                                           const double* __restrict__ X, uint32_t n, double* Y) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  double s[M];
#pragma unroll
  for (int c = 0; c < M; ++c) s[c] = 0.0;
  const uint32_t e1 = colPtr[j + 1];
  for (uint32_t e = colPtr[j]; e < e1; e += kSpmmUnroll) {
    uint32_t i[kSpmmUnroll];
    double v[kSpmmUnroll];
#pragma unroll
    for (int u = 0; u < kSpmmUnroll; ++u) {
      const uint32_t eu = min(e + u, e1 - 1);
      i[u] = rows[eu];
      v[u] = vals[eu];
    }
    double x[kSpmmUnroll][M];
#pragma unroll
    for (int u = 0; u < kSpmmUnroll; ++u)
#pragma unroll
      for (int c = 0; c < M; ++c) x[u][c] = X[uint64_t(i[u]) * M + c];
#pragma unroll
    for (int u = 0; u < kSpmmUnroll; ++u) {
      const bool on = e + u < e1;
#pragma unroll
      for (int c = 0; c < M; ++c) {
        const double prod = x[u][c] * v[u];
        s[c] = on ? s[c] + prod : s[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < M; ++c) Y[uint64_t(j) * M + c] = s[c];
}

// Block partials of the dot products a . b of columns of A and B: all pairs (SYM = 0) or a <= b
// (SYM = 1, A = B); part[pair * gridDim.x + block]
template <int M, int SYM>
static __global__ __launch_bounds__(kBlock) void dotPartialKernel(const double* A, const double* B, uint32_t n,
                                                                 double* part) {
  constexpr int P = SYM ? M * (M + 1) / 2 : M * M;
  double acc[P];
#pragma unroll
  for (int q = 0; q < P; ++q) acc[q] = 0.0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double a[M], b[M];
#pragma unroll
    for (int c = 0; c < M; ++c) {
      a[c] = A[uint64_t(i) * M + c];
      b[c] = SYM ? a[c] : B[uint64_t(i) * M + c];
    }
    int q = 0;
#pragma unroll
    for (int x = 0; x < M; ++x)
#pragma unroll
      for (int y = SYM ? x : 0; y < M; ++y, ++q) {
        const double prod = a[x] * b[y];
        acc[q] += prod;
      }
  }
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const double v = blockReduceSumF64(acc[q]);
    if (threadIdx.x == 0) part[uint64_t(q) * gridDim.x + blockIdx.x] = v;
  }
}

// Sums the partials of one pair in block order (threads stride the blocks, then a fixed tree)
__device__ inline double sumPartials(const double* p, uint32_t nb) {
  double t = 0.0;
  for (uint32_t q = threadIdx.x; q < nb; q += kBlock) t += p[q];
  return blockReduceSumF64(t);
}

// Gram matrix of Y (upper triangle from dotPartialKernel, SYM = 1) and its Cholesky factor
// G = R^T R (R upper, row-major m x m); *bad is set (and stays set) if G is not positive definite.
template <int M>
static __global__ __launch_bounds__(kBlock) void cholKernel(const double* part, uint32_t nb, double* R, int* bad) {
  __shared__ double G[M * M];
  int pair = 0;
  for (int a = 0; a < M; ++a)
    for (int b = a; b < M; ++b, ++pair) {
      const double g = sumPartials(part + uint64_t(pair) * nb, nb);
      if (threadIdx.x == 0) G[a * M + b] = g;
    }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double r[M * M];
#pragma unroll
  for (int q = 0; q < M * M; ++q) r[q] = 0.0;
  for (int j = 0; j < M; ++j) {
    double d = G[j * M + j];
    for (int k = 0; k < j; ++k) d -= r[k * M + j] * r[k * M + j];
    if (!(d > 0.0)) {
      *bad = 1;
      d = 1.0;
    }
    r[j * M + j] = sqrt(d);
    for (int l = j + 1; l < M; ++l) {
      double v = G[j * M + l];
      for (int k = 0; k < j; ++k) v -= r[k * M + j] * r[k * M + l];
      r[j * M + l] = v / r[j * M + j];
    }
  }
  for (int q = 0; q < M * M; ++q) R[q] = r[q];
}

// In place: x_c = (y_c - sum_{d<c} R_dc x_d) / R_cc, d ascending
template <int M>
static __global__ __launch_bounds__(kBlock) void applyRinvKernel(double* Y, uint32_t n, const double* R) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x[M];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    double v = Y[uint64_t(i) * M + c];
#pragma unroll
    for (int d = 0; d < c; ++d) v -= R[d * M + c] * x[d];
    x[c] = v / R[c * M + c];
  }
#pragma unroll
  for (int c = 0; c < M; ++c) Y[uint64_t(i) * M + c] = x[c];
}

// H_ab = x_a . z_b from dotPartialKernel (SYM = 0): one block
static __global__ __launch_bounds__(kBlock) void ritzKernel(const double* part, uint32_t nb, int m, double* H) {
  for (int q = 0; q < m * m; ++q) {
    const double h = sumPartials(part + uint64_t(q) * nb, nb);
    if (threadIdx.x == 0) H[q] = h;
  }
}

// box-by-box basis to column-by-column (the C ABI's layout)
static __global__ __launch_bounds__(kBlock) void transposeBasisKernel(const double* X, uint32_t n, int m,
                                                                     double* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < m; ++c) out[uint64_t(c) * n + i] = X[uint64_t(i) * m + c];
}

// Plain copy (traffic-counter calibration): every byte read once and written once.
static __global__ __launch_bounds__(kBlock) void copyKernel(const uint4* in, uint4* out, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock)
    out[i] = in[i];
}

static __global__ __launch_bounds__(kBlock) void piFinishKernel(const double* y, const double* mu0, uint32_t n,
                                                         const PiState* st, double* mu) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t f = st->final;
  mu[j] = f == 0 ? mu0[j] : y[uint64_t(f % 3) * n + j] / st->norm[f % 3];
}

}  // namespace b12x
