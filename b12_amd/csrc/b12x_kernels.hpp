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

// Exclusive scan over a block of BS threads; returns the exclusive prefix of v and the block total.
template <int BS = kBlock>
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t waveSums[BS / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t inc = waveInclusiveScan(v);
  if (lane == kWave - 1) waveSums[wave] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BS / kWave; ++w) {
    const uint32_t s = waveSums[w];
    before += (w < wave) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

// *flag = 1 when some lane of the wave has `set`: one atomic per wave, none once the flag reads as
// set (same-address atomics serialise chip-wide). Lane 0 of the wave must call it; lanes that have
// left count as not set.
__device__ inline void setFlagOncePerWave(bool set, uint32_t* flag) {
  if (__ballot(set) && (threadIdx.x & (kWave - 1)) == 0 &&
      __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
    atomicOr(flag, 1u);
}

template <int BS = kBlock>
__device__ inline uint32_t blockReduceSum(uint32_t v) {
  __shared__ uint32_t waveSums[BS / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  v = waveReduceSum(v);
  if (lane == 0) waveSums[wave] = v;
  __syncthreads();
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < BS / kWave; ++w) tot += waveSums[w];
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

// The end of a subdivision step replayed from a graph: its counts scal[0..3] are appended to a
// pinned host log (entry *index, which advances), and when the step kept no leaf the surrounding
// box is written as the only leaf (ImplicitBoxTree.hpp:1959-1970) and *kept becomes 1, so the
// next step reads a count of 1.
struct StepLogArgs {
  uint32_t* index;
  uint32_t* hostLog;
  uint64_t* key;  // the leaf arrays the step wrote
  uint8_t* depth;
  uint8_t* flags;
  uint64_t rootKey;
};

__device__ inline void logStep(const StepLogArgs& g, const uint32_t* scal, uint32_t* kept) {
  const uint32_t i = *g.index;
  *g.index = i + 1;
#pragma unroll
  for (int j = 0; j < 4; ++j) g.hostLog[4 * i + j] = scal[j];
  if (*kept == 0) {
    g.key[0] = g.rootKey;
    g.depth[0] = 0;
    g.flags[0] = 0;
    *kept = 1;
  }
}

constexpr int kSmallBlock = 1024;  // threads of the one-block step kernels

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

// The end of a captured small subdivision step in one block, in place: the keep of step t
// (remove(HIT) with the hit bytes counting as HIT, then unsetFlags(HIT): ImplicitBoxTree.hpp:
// 2343-2371, 3334-3411) and, while the batch has a step left, the subdivide(sd) of step t + 1
// (:2883-2930) in the same pass: a kept leaf emits its two halves, or itself when it lacks sd or
// lies at the maximum depth, at its scan position. With hit == nullptr it is the batch's first
// subdivide alone (every leaf stays with its flags). Every thread reads all its leaves before the
// block's scan barriers and writes only after them, so the output overwrites the input safely.
// At most kKeepSubMax leaves in and out. Counts: scal[0] leaves after the subdivide, scal[1]
// leaves with sd, scal[2] leaves kept, scal[3] leaves newly hit. A keep appends its step's entry
// (scal[0], scal[1], kept, newly hit) to the pinned log before scal[0..1] move on to the next
// step, and when it kept no leaf the surrounding box takes their place (ImplicitBoxTree.hpp:
// 1959-1970), subdivided like any kept leaf.
constexpr int kKeepSubItems = 16;
constexpr uint32_t kKeepSubMax = uint32_t(kSmallBlock) * kKeepSubItems;

struct KeepSubArgs {
  TreeView v;           // the leaves, read and rewritten in place
  uint8_t* hit;         // the step's hit bytes (cleared here); nullptr: the batch's first subdivide
  uint32_t* scal;
  uint32_t* stepsLeft;  // subdivides left in the batch: a keep takes one; the first subdivide sets it
  uint32_t stepsInit;
  uint32_t* logIndex;   // entries in the pinned log: a keep appends one; the first subdivide resets it
  uint32_t* hostLog;
  uint8_t sd;
  uint64_t rootKey;
};

// the leaf (k, d, f) as c = 1 or 2 leaves at pos (two: its halves)
__device__ inline void emitLeaf(const TreeView& v, uint32_t pos, uint64_t k, int d, uint8_t f, uint32_t c) {
  if (c == 1) {
    v.key[pos] = k;
    v.depth[pos] = uint8_t(d);
    v.flags[pos] = f;
  } else {
    const uint64_t b = bitsOf(k, d) << 1;
    v.key[pos] = keyOf(b, d + 1);
    v.key[pos + 1] = keyOf(b | 1ull, d + 1);
    v.depth[pos] = v.depth[pos + 1] = uint8_t(d + 1);
    v.flags[pos] = v.flags[pos + 1] = f;
  }
}

static __global__ __launch_bounds__(kSmallBlock) void keepSubdivideKernel(KeepSubArgs a) {
  const bool keepPass = a.hit != nullptr;
  const uint32_t n = min(keepPass ? a.scal[0] : a.scal[2], kKeepSubMax);
  const bool doSub = keepPass ? (*a.stepsLeft > 0) : true;
  const uint32_t base = threadIdx.x * uint32_t(kKeepSubItems);
  uint64_t key[kKeepSubItems];
  uint32_t packed[kKeepSubItems];  // depth | flags << 8 | count << 16
  uint32_t s = 0, kept = 0, fresh = 0, withSd = 0;
#pragma unroll
  for (int t = 0; t < kKeepSubItems; ++t) {
    const uint32_t i = base + uint32_t(t);
    key[t] = 0;
    packed[t] = 0;
    if (i < n) {
      key[t] = a.v.key[i];
      const uint32_t d = a.v.depth[i];
      uint8_t f = a.v.flags[i];
      bool keep = true;
      if (keepPass) {
        const bool h = a.hit[i] != 0;
        fresh += (h && !areFlagsSet(f, F_HIT)) ? 1u : 0u;
        keep = h || areFlagsSet(f, F_HIT);
        f = uint8_t(f & ~uint8_t(F_HIT));
      }
      uint32_t c = 0;
      if (keep) {
        ++kept;
        const bool sdSet = doSub && areFlagsSet(f, a.sd);
        withSd += sdSet ? 1u : 0u;
        c = (sdSet && d < uint32_t(kMaxDepth)) ? 2u : 1u;
      }
      s += c;
      packed[t] = d | (uint32_t(f) << 8) | (c << 16);
    }
  }
  uint32_t nOut = 0;
  uint32_t pos = blockExclusiveScan<kSmallBlock>(s, nOut);
  kept = blockReduceSum<kSmallBlock>(kept);
  fresh = blockReduceSum<kSmallBlock>(fresh);
  withSd = blockReduceSum<kSmallBlock>(withSd);
  // every leaf of the block has been read (the barriers above): write in place
#pragma unroll
  for (int t = 0; t < kKeepSubItems; ++t) {
    const uint32_t c = packed[t] >> 16;
    if (c) emitLeaf(a.v, pos, key[t], int(packed[t] & 0xFFu), uint8_t((packed[t] >> 8) & 0xFFu), c);
    pos += c;
    if (keepPass && base + uint32_t(t) < n) a.hit[base + t] = 0;
  }
  if (threadIdx.x == 0) {
    if (kept == 0) {  // only a keep can empty the tree: the surrounding box, with no flags
      const bool sdSet = doSub && areFlagsSet(uint8_t(0), a.sd);
      const uint32_t c = sdSet ? 2u : 1u;
      emitLeaf(a.v, 0, a.rootKey, 0, 0, c);
      nOut = c;
      withSd = sdSet ? 1u : 0u;
    }
    if (keepPass) {
      const uint32_t li = *a.logIndex;
      *a.logIndex = li + 1;
      a.hostLog[4 * li + 0] = a.scal[0];
      a.hostLog[4 * li + 1] = a.scal[1];
      a.hostLog[4 * li + 2] = kept;
      a.hostLog[4 * li + 3] = fresh;
      a.scal[2] = kept ? kept : 1u;
      a.scal[3] = fresh;
      if (doSub) {
        a.scal[0] = nOut;
        a.scal[1] = withSd;
        *a.stepsLeft -= 1;
      }
    } else {
      a.scal[0] = nOut;
      a.scal[1] = withSd;
      *a.stepsLeft = a.stepsInit;
      *a.logIndex = 0;
    }
  }
}

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
  else if (n > 0) {
    // locate() with a trip count that depends on n only (wave-uniform probes, no branches)
    const uint64_t pk1[1] = {pk};
    uint32_t t1[1];
    locateN<1>(a.key, a.depth, n, pk1, t1);
    t = t1[0];
  } else {
    t = kInvalid;
  }
  if (a.outKey) a.outKey[i] = pk;
  if (a.mode == 0) {
    if (t != kInvalid) a.hit[t] = 1;
  } else {
    a.out[i] = t;
  }
}

// PPL points of one lane at once (local indices idx[q], idx[0] < np): lanes past the end repeat
// point idx[0] and store nothing. Same arithmetic per point as imagePoint. SAMEBOX: the PPL
// points are consecutive points of one box (the caller checks that k and the first global index
// are multiples of PPL), so the box's geometry and index arithmetic are formed once.
template <int DIM, int MAP, int GRID, int PPL, bool SAMEBOX = false>
__device__ inline void imagePointsN(const ImageArgs& a, const uint32_t (&idx)[PPL], uint32_t np, uint32_t n) {
  const uint32_t k = a.grid.k, base = uint32_t(a.pointBase);
  double x[PPL][DIM], y[PPL][DIM];
  bool valid[PPL];
  if constexpr (SAMEBOX) {
    const uint32_t gi0 = base + idx[0];
    const uint32_t b = gi0 / k, gl0 = gi0 - b * k;
    const uint32_t src = a.srcList ? a.srcList[b] : b;
    double c[DIM], rb[DIM];
    boxCoords<DIM>(a.dom, a.srcKey[src], a.srcDepth[src], c, rb);
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      valid[q] = idx[q] < np;
      double g[DIM];
      if constexpr (GRID == GRID_INNER) {
#pragma unroll
        for (int j = 0; j < DIM; ++j) g[j] = a.grid.table[(gl0 + q) * DIM + j];
      } else {
        uniformGridPoint<DIM>(a.grid.seed, gi0 + q, g);
      }
#pragma unroll
      for (int j = 0; j < DIM; ++j) x[q][j] = c[j] + g[j] * rb[j];
    }
  } else {
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

// Any block size up to kBlock (the launch picks it); PPL points per lane. SAMEBOX: a lane's PPL
// points are consecutive points of one box (when k and the first global index are multiples of
// PPL; otherwise points BS apart, as without SAMEBOX).
template <int DIM, int MAP, int GRID, int PPL, bool SAMEBOX = false>
static __global__ __launch_bounds__(kBlock) void imageKernel(ImageArgs a) {
  const uint32_t BS = blockDim.x;
  __shared__ uint64_t sm[PPL == 1 ? kDescentSamples : 1];
  // a device-side count is capped by the host's bound (a prefix when the launch is a sample)
  const uint64_t nPoints = a.nBoxesDev ? min(uint64_t(*a.nBoxesDev) * a.grid.k, a.nPoints) : a.nPoints;
  const uint32_t n = a.nBoxesDev ? *a.nBoxesDev : a.n;
  if (uint64_t(blockIdx.x) * BS * PPL >= nPoints) return;  // grid sized for a bound: no points here
  if constexpr (PPL > 1 && SAMEBOX) {
    if (a.grid.k % PPL == 0 && a.pointBase % PPL == 0 && !a.depth0 &&
        a.pointBase + nPoints + uint64_t(PPL) * BS <= 0xFFFFFFFFull &&
        uint64_t(gridDim.x) * BS * PPL <= 0xFFFFFFFFull) {
      const uint32_t np = uint32_t(nPoints), stride = gridDim.x * BS * PPL;
      for (uint32_t i0 = (blockIdx.x * BS + threadIdx.x) * PPL; i0 < np; i0 += stride) {
        uint32_t idx[PPL];
#pragma unroll
        for (int q = 0; q < PPL; ++q) idx[q] = i0 + q;
        imagePointsN<DIM, MAP, GRID, PPL, true>(a, idx, np, n);
      }
      return;
    }
  }
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
  // lanes past P have returned; lane 0 holds the wave's smallest index, so it is here if any is
  if (anyValid) setFlagOncePerWave(valid, anyValid);
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
  __shared__ int blockSawInvalid;
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / S, pos = lane % S;
  if (threadIdx.x == 0) blockSawInvalid = 0;
  __syncthreads();
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
    if (a.pass == 0 && __ballot(invalidHere) && lane == 0) blockSawInvalid = 1;
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
  // at most one atomic per block, and none once the flag reads as set (same-address atomics
  // serialise chip-wide: one per wave took 70 ms for 18 M rows)
  __syncthreads();
  if (threadIdx.x == 0 && blockSawInvalid &&
      __hip_atomic_load(a.anyInvalid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
    atomicOr(a.anyInvalid, 1u);
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

// row[i] = rowBase + i / k: the source box of point i of a pass starting at box rowBase
static __global__ __launch_bounds__(kBlock) void rowOfPointKernel(uint32_t* row, uint64_t n, uint32_t k,
                                                                  uint32_t rowBase) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) row[i] = rowBase + uint32_t(i / k);
}

static __global__ __launch_bounds__(kBlock) void anyInvalidKernel(const uint32_t* col, uint64_t n, uint32_t* flag) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  setFlagOncePerWave(i < n && col[i] == 0xFFFFFFFFu, flag);
}

// Expand CSR row pointers into the COO row index per entry (rowBase + the local row).
static __global__ __launch_bounds__(kBlock) void expandRowsKernel(const uint32_t* rowPtr, uint64_t nRows,
                                                                  uint32_t rowBase, uint32_t* rows) {
  const uint64_t r = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= nRows) return;
  for (uint32_t e = rowPtr[r]; e < rowPtr[r + 1]; ++e) rows[e] = rowBase + uint32_t(r);
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

// Block sum in a fixed order: each wave by a butterfly (every lane ends with the same value,
// since a + b = b + a in IEEE arithmetic), then the wave sums added to 0.0 in wave order. For a
// block of kBlock lanes this is the sum of one 256-column chunk of the power iteration's order.
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

// ---------------------------------------------------------------------------------------------
// Invariant measure: power iteration on unnormalised iterates (build-defined; the reference leaves
// this step to MATLAB, subdivisionByNim.m:30-59), row-vector convention:
//   z_0 = mu_0, nu_0 = 1;  z_t[j] = c_t * sum_{i in column j, ascending} z_{t-1}[i] * P_ij;
//   nu_t = S(|z_t|);  r_t = S(|z_t[j] / nu_t - z_{t-1}[j] / nu_{t-1}|);
//   stop at the first t with (tol >= 0 and r_t <= tol) or t = maxIter;  mu = z_t / nu_t.
// c_t (piScale) is 1 except at t = 16, 24, ... where it is 2^-ilogb(nu_{t-8}) when that exponent
// leaves [-32, 32]: a power of two, so z_t and nu_t scale exactly and mu_t is that of the
// normalised iteration with the normalisation deferred. No iterate depends on a norm of the
// iteration before it, which is what lets the persistent form below run without a grid-wide
// barrier per iteration.
// S sums in one fixed order: a chunk of 256 columns is four groups of 64, each group summed by a
// wave butterfly (x_l += x_{l^o}, o = 32 ... 1; columns past n are 0) and the four group sums
// added to 0.0 in order; lane l of one wave adds chunks l, l + 64, ... to 0.0 in ascending order
// and the lanes combine by the butterfly. The persistent and batched forms here, the forest's
// pools and the oracle all use exactly these operations, so their iterates, norms, stopping
// decisions and measures agree bit for bit.
// ---------------------------------------------------------------------------------------------
struct PiState {
  double norm[3];   // batched form: nu_t in slot t % 3
  double resid;     // residual of the last completed iteration
  double nuLag;     // batched form: nu_t of the last t with t % 8 == 0 (c_{t+8} derives from it)
  double nuFinal;   // nu of the final iteration
  uint32_t t;       // batched form: the next iteration
  int32_t done;
  uint32_t final;   // the final iteration (0: mu = mu_0)
  uint32_t slot;    // ring slot holding z_final
};

__device__ inline double piScale(uint32_t t, double nuLag) {
  if (t < 16 || (t & 7u) != 0) return 1.0;
  if (!(nuLag > 0.0) || !isfinite(nuLag)) return 1.0;
  const int e = ilogb(nuLag);
  if (e >= -32 && e <= 32) return 1.0;
  return ldexp(1.0, -max(-1000, min(1000, e)));
}

__device__ inline double waveButterflySum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Batched form, one launch per iteration t: z_t for kBlock columns per block (one lane per
// column, kPiUnroll entries' loads in flight), with the chunk sums of |z_t| and of the residual
// terms of iteration t - 1 (the block is one chunk: blockReduceSumF64 is the chunk order).
// z rotates over 3 slots (z_t in slot t % 3).
constexpr int kPiUnroll = 8;

static __global__ __launch_bounds__(kBlock) void piStepKernel(const uint32_t* __restrict__ colPtr,
                                                       const uint32_t* __restrict__ rowIdx,
                                                       const double* __restrict__ val, const double* __restrict__ mu0,
                                                       double* z, uint32_t n, double* partAbs, double* partRes,
                                                       const PiState* st) {
  if (st->done) return;
  const uint32_t t = st->t;
  const double* zPrev = t == 1 ? mu0 : z + uint64_t((t + 2) % 3) * n;   // z_{t-1}
  const double* zPrev2 = t == 2 ? mu0 : z + uint64_t((t + 1) % 3) * n;  // z_{t-2} (t >= 2)
  double* zCur = z + uint64_t(t % 3) * n;
  const double c = piScale(t, st->nuLag);
  const double nu1 = t >= 2 ? st->norm[(t + 2) % 3] : 1.0;
  const double nu2 = t >= 3 ? st->norm[(t + 1) % 3] : 1.0;
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  double sAbs = 0.0, sRes = 0.0;
  if (j < n) {
    double s = 0.0;
    const uint32_t e1 = colPtr[j + 1];
    for (uint32_t e = colPtr[j]; e < e1; e += kPiUnroll) {
      uint32_t i[kPiUnroll];
      double v[kPiUnroll], m[kPiUnroll];
#pragma unroll
      for (int u = 0; u < kPiUnroll; ++u) {
        const uint32_t eu = min(e + u, e1 - 1);
        i[u] = rowIdx[eu];
        v[u] = val[eu];
      }
#pragma unroll
      for (int u = 0; u < kPiUnroll; ++u) m[u] = zPrev[i[u]];
#pragma unroll
      for (int u = 0; u < kPiUnroll; ++u) {
        const double prod = m[u] * v[u];
        s = (e + u < e1) ? s + prod : s;
      }
    }
    if (c != 1.0) s *= c;
    zCur[j] = s;
    sAbs = fabs(s);
    if (t >= 2) sRes = fabs(zPrev[j] / nu1 - zPrev2[j] / nu2);
  }
  sAbs = blockReduceSumF64(sAbs);
  sRes = blockReduceSumF64(sRes);
  if (threadIdx.x == 0) {
    partAbs[blockIdx.x] = sAbs;
    partRes[blockIdx.x] = sRes;
  }
}

// One wave after every piStepKernel: sums the chunk partials (lane l: chunks l, l + 64, ... in
// order, then the butterfly), records nu_t, and decides whether iteration t - 1 ends the loop.
static __global__ __launch_bounds__(kWave) void piReduceKernel(const double* partAbs, const double* partRes, uint32_t nb,
                                                        PiState* st, uint32_t maxIter, double tol) {
  if (st->done) return;
  const uint32_t t = st->t, lane = threadIdx.x;
  constexpr int U = 8;  // loads in flight per lane before they are added (order unchanged)
  double ta = 0.0, tr = 0.0;
  for (uint32_t q0 = 0; q0 < nb; q0 += kWave * U) {
    double va[U], vr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t q = q0 + u * kWave + lane;
      va[u] = q < nb ? partAbs[q] : 0.0;
      vr[u] = q < nb ? partRes[q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (q0 + u * kWave + lane < nb) {
        ta += va[u];
        tr += vr[u];
      }
  }
  ta = waveButterflySum(ta);
  tr = waveButterflySum(tr);
  if (lane == 0) {
    st->norm[t % 3] = ta;
    if ((t & 7u) == 0) st->nuLag = ta;
    if (t >= 2) {
      st->resid = tr;
      if ((tol >= 0.0 && tr <= tol) || t - 1 >= maxIter) {
        st->done = 1;
        st->final = t - 1;
        st->nuFinal = st->norm[(t + 2) % 3];
        st->slot = (t - 1) % 3;
      }
    }
    st->t = t + 1;
  }
}

// ---------------------------------------------------------------------------------------------
// Persistent form (piFlowKernel): one launch for the whole solve, its grid resident at once.
// Worker waves own groups of 64 columns (wave w of block b: groups 4b + w, 4b + w + 4 G, ...
// for G worker blocks), one lane per column; the last block is the aggregator.
// Iteration t of a worker wave gathers z_{t-1} from the slots of the other waves, stores z_t and
// publishes its group's sums of |z_t| and of the residual terms of iteration t - L (L =
// kFlowLag). It waits only for the entries it gathers (dataflow through the vector: a ring of
// kFlowRing slots, each entry the empty pattern until its owner stores it) and for the totals
// of iteration t - L, which the aggregator has published long before in the steady state. So
// no grid-wide barrier lies on the path of an iteration. In the straight-line form the totals
// load is issued together with the gathers, and an iteration's group sums are formed and
// published in the next iteration, while its gathers are in flight.
// The aggregator sums the group partials of iteration tau in the chunk order and publishes
// (nu_tau, r_{tau - L}) in totals slot tau % kFlowTot (kFlowCopies copies, one 128-byte line
// each; wave w of block b polls copy (4 b + w) % kFlowCopies), two iterations at a time. The stopping decision for
// iteration f = tau - L is taken once r_f is in the totals: the aggregator records it and
// leaves, and every worker wave reads the same totals at its iteration f + 2 L and leaves
// there, before z_f is cleared (iteration f + kFlowClear), so piFinishKernel forms z_f / nu_f.
// Slot reuse (D = kFlowClear; the static_asserts below hold the bounds):
//  - z: in iteration t a lane stores z_t in slot t % kFlowRing and clears z_{t-D}: its readers
//    gathered it in iteration t - D + 1, and having seen the totals of t - L the lane knows every
//    wave has completed the gathers of iteration t - L >= t - D + 1. A reader gathering z_x from
//    the slot has seen z_{x-1} of the same lane, which was stored after the lane's store counter
//    drained past the clear (issued in iteration x - kFlowRing + D <= x - 2), so it finds empty
//    or z_x.
//  - partials: the aggregator takes iterations in pairs (tau, tau + 1) and clears their positions
//    after publishing the pair's totals; the clears drain before the next pair's totals are
//    published, and a wave writes the partials of tau + 8 (tau + 9) only after seeing the totals
//    of tau + 2 (tau + 3).
//  - totals: after publishing the pair it also clears the slots of the totals of tau - 6 and
//    tau - 5 (every wave read them before it published the partials of tau and tau + 1); their
//    next totals, tau + 10 and tau + 11, are published five pairs later, and a wave reading them
//    has seen those of tau + 9, published after the clears drained.
// Everything written during the run goes through agent-scope relaxed atomics, coherent across
// XCDs without L2 write-back or invalidation; the matrix is read-only and stays cached. Every
// wait is bounded: past spinLimit polls a wave raises *err and leaves, the others leave at their
// next poll, and the host recomputes with the batched form.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kFlowMaxWorkers = 512;
constexpr uint32_t kFlowRing = 16, kFlowLag = 6, kFlowClear = 2 * kFlowLag, kFlowPart = 8, kFlowTot = 16;
static_assert(kFlowClear >= 2 * kFlowLag, "z_f must outlive the workers");
static_assert(kFlowClear + 2 <= kFlowRing, "a clear drains before the slot's next value");
static_assert(kFlowLag + 1 <= kFlowClear, "readers of a cleared value have finished");
static_assert(kFlowPart >= kFlowLag + 2, "partial reuse: positions of a pair are rewritten after the next pair");
static_assert(kFlowTot >= kFlowLag + 10, "totals reuse: a cleared slot's next totals come five pairs later");
static_assert(kFlowLag <= 8, "c_t derives from nu_{t-8}, which the totals history holds");
constexpr uint32_t kFlowCopies = 16, kFlowStride = 16;  // u64 words per totals copy: one 128-byte line
constexpr uint32_t kFlowTotWords = kFlowTot * kFlowCopies * kFlowStride;
constexpr int kCoopUnroll = 8;
constexpr int kCoopCache = 3 * kCoopUnroll;  // entries of a lane's first column cached for the whole run
constexpr uint64_t kCoopEmpty = 0xFFF4B12B12B12B12ull;  // a NaN pattern arithmetic never produces
constexpr uint32_t kFlowAggChunks = 4;  // chunks per aggregator lane polled together

struct PiFlowArgs {
  const uint32_t* colPtr;
  const uint32_t* rowIdx;
  const double* val;
  const double* mu0;
  uint64_t* z;     // kFlowRing * n, all kCoopEmpty at launch
  uint64_t* part;  // kFlowPart * nGroups * 2 (|z| sum, residual sum), all kCoopEmpty at launch
  uint64_t* tot;   // kFlowTotWords, all kCoopEmpty at launch
  PiState* st;
  int* err;        // zero at launch
  uint32_t n, nGroups, workers;
  uint32_t maxIter;
  double tol;
  uint32_t spinLimit;
  const uint32_t* maxLen;  // longest column (the straight-line form checks it against its cache)
};

constexpr int kFlowNeedsSweep = 2;  // *err when a column is longer than the straight-line form caches

template <class T>
__device__ inline T ldAgent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void stAgent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double asF64(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ inline uint64_t asU64(double v) { return uint64_t(__double_as_longlong(v)); }

// entry at a byte offset below 2^32 from a uniform base (one 32-bit offset register per gather)
__device__ inline uint64_t ldAt(const uint64_t* base, uint32_t off) {
  return ldAgent(reinterpret_cast<const uint64_t*>(reinterpret_cast<const char*>(base) + off));
}

// Issues the gathers m[u] = src[off u] of a lane's entries u < cnt (byte offsets), all in flight
// together; a lane loads only its own entries (a group of kCoopUnroll entries that no lane of the
// wave has is skipped by one wave-uniform test). Loads for lanes past their entries measured
// slower on MI355X whether they repeated a lane's last entry or read a shared or own line: every
// extra address costs the memory pipeline and the lines they hit are being written.
template <int U>
__device__ inline void flowIssue(const uint64_t* src, const uint32_t (&off)[U], uint32_t cnt, uint64_t (&m)[U]) {
#pragma unroll
  for (int g = 0; g < U; g += kCoopUnroll) {
    if (__any(uint32_t(g) < cnt)) {
#pragma unroll
      for (int u = g; u < g + kCoopUnroll; ++u) m[u] = uint32_t(u) < cnt ? ldAt(src, off[u]) : 0;
    } else {
#pragma unroll
      for (int u = g; u < g + kCoopUnroll; ++u) m[u] = 0;
    }
  }
}

// Completes flowIssue: re-reads the entries that came back empty until none is left (check: the
// source is a ring slot still being filled). false when a wait passed its limit or another wave
// raised the error flag.
template <int U>
__device__ inline bool flowComplete(const uint64_t* src, const uint32_t (&off)[U], uint32_t cnt, bool check,
                                    uint64_t (&m)[U], const PiFlowArgs& a) {
  if (!check) return true;
  for (uint32_t spin = 0;; ++spin) {
    bool miss = false;
#pragma unroll
    for (int u = 0; u < U; ++u) miss = miss || (uint32_t(u) < cnt && m[u] == kCoopEmpty);
    if (!__any(miss)) return true;
    if (spin >= a.spinLimit || ldAgent(a.err)) return false;
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int g = 0; g < U; g += kCoopUnroll) {
      bool gm = false;
#pragma unroll
      for (int u = g; u < g + kCoopUnroll; ++u) gm = gm || (uint32_t(u) < cnt && m[u] == kCoopEmpty);
      if (__any(gm)) {
#pragma unroll
        for (int u = g; u < g + kCoopUnroll; ++u)
          if (uint32_t(u) < cnt && m[u] == kCoopEmpty) m[u] = ldAt(src, off[u]);
      }
    }
  }
}

__device__ inline void flowAbort(const PiFlowArgs& a) { stAgent(a.err, 1); }

// sum over u of z_u v_u for the lane's cached entries in ascending order, times c when c != 1:
// the groups of kCoopUnroll slots beyond every lane's column are skipped (wave-uniform), the
// others added whole (a slot past the column adds -0.0, which leaves the sum as it is)
__device__ inline double flowCachedSum(const uint64_t (&m)[kCoopCache], const double (&cv)[kCoopCache][kBlock],
                                       uint32_t cnt, double c) {
  double s = 0.0;
#pragma unroll
  for (int g = 0; g < kCoopCache; g += kCoopUnroll) {
    if (__any(uint32_t(g) < cnt)) {
#pragma unroll
      for (int u = g; u < g + kCoopUnroll; ++u) s += asF64(m[u]) * cv[u][threadIdx.x];
    }
  }
  return c != 1.0 ? s * c : s;
}

// The aggregator takes iterations two at a time (tau, tau + 1): it polls both iterations' group
// partials together (one loaded round trip for the pair), sums each in the chunk order, drains its
// earlier clears, publishes (nu_tau, r_{tau-L}) and then tau + 1's, and only then clears the two
// consumed partial positions and the slots of the totals of tau - 6 and tau - 5 (the reuse rules
// above). The clears drain while the next pair's loads are in flight. The stopping decision is
// taken in iteration order; the totals a worker needs for iteration t (those of t - L) depend on
// partials of iterations up to t - L + 1, which every worker has published by iteration t.
__device__ inline void piFlowAggregator(const PiFlowArgs& a) {
  __shared__ double nuRing[kFlowTot];
  const uint32_t lane = threadIdx.x;
  const uint32_t nG = a.nGroups, nChunks = (nG + 3) / 4;
  for (uint32_t tau0 = 1;; tau0 += 2) {
    double ta[2] = {0.0, 0.0}, tr[2] = {0.0, 0.0};
    // lane l sums chunks l, l + 64, ... in order; a chunk's sum adds its four group partials to
    // 0.0 in order. kFlowAggChunks chunks per lane are polled together, for both iterations.
    for (uint32_t k0 = 0; k0 * kWave < nChunks; k0 += kFlowAggChunks) {
      uint64_t v[2][kFlowAggChunks][8];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (uint32_t k = 0; k < kFlowAggChunks; ++k)
#pragma unroll
          for (int x = 0; x < 8; ++x) v[p][k][x] = 0;
      for (uint32_t spin = 0;; ++spin) {
        bool miss = false;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const uint64_t* buf = a.part + uint64_t((tau0 + p) % kFlowPart) * nG * 2;
#pragma unroll
          for (uint32_t k = 0; k < kFlowAggChunks; ++k) {
            const uint32_t q = lane + (k0 + k) * kWave;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
              const uint32_t g = 4 * q + uint32_t(x / 2);
              if (q < nChunks && g < nG) {
                if (spin == 0 || v[p][k][x] == kCoopEmpty) v[p][k][x] = ldAgent(buf + uint64_t(g) * 2 + (x & 1));
              }
            }
          }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (uint32_t k = 0; k < kFlowAggChunks; ++k) {
            const uint32_t q = lane + (k0 + k) * kWave;
#pragma unroll
            for (int x = 0; x < 8; ++x)
              miss = miss || (q < nChunks && 4 * q + uint32_t(x / 2) < nG && v[p][k][x] == kCoopEmpty);
          }
        if (!__any(miss)) break;
        if (spin >= a.spinLimit || ldAgent(a.err)) {
          if (lane == 0) flowAbort(a);
          return;
        }
        __builtin_amdgcn_s_sleep(1);
      }
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (uint32_t k = 0; k < kFlowAggChunks; ++k) {
          const uint32_t q = lane + (k0 + k) * kWave;
          if (q < nChunks) {
            double ca = 0.0, cr = 0.0;
#pragma unroll
            for (int w = 0; w < 4; ++w)
              if (4 * q + uint32_t(w) < nG) {
                ca += asF64(v[p][k][2 * w]);
                cr += asF64(v[p][k][2 * w + 1]);
              }
            ta[p] += ca;
            tr[p] += cr;
          }
        }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      ta[p] = waveButterflySum(ta[p]);
      tr[p] = waveButterflySum(tr[p]);
    }
    __builtin_amdgcn_s_waitcnt(0);  // the clears of the previous pair have completed
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const uint32_t tau = tau0 + p;
      if (lane == 0) nuRing[tau % kFlowTot] = ta[p];
      if (lane < kFlowCopies) {
        uint64_t* tc = a.tot + (uint64_t(tau % kFlowTot) * kFlowCopies + lane) * kFlowStride;
        stAgent(tc, asU64(ta[p]));
        stAgent(tc + 1, asU64(tr[p]));
      }
      if (tau > kFlowLag) {
        const uint32_t f = tau - kFlowLag;
        if ((a.tol >= 0.0 && tr[p] <= a.tol) || f >= a.maxIter) {
          __builtin_amdgcn_wave_barrier();
          if (lane == 0) {
            a.st->final = f;
            a.st->resid = tr[p];
            a.st->nuFinal = nuRing[f % kFlowTot];
            a.st->slot = f % kFlowRing;
            a.st->done = 1;
          }
          return;
        }
      }
    }
    // clear the consumed partials of tau0 and tau0 + 1 and the totals of tau0 - 6 and tau0 - 5
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      uint64_t* pc = a.part + uint64_t((tau0 + p) % kFlowPart) * nG * 2;
      for (uint32_t g = lane; g < nG; g += kWave) {
        stAgent(pc + uint64_t(g) * 2, kCoopEmpty);
        stAgent(pc + uint64_t(g) * 2 + 1, kCoopEmpty);
      }
      if (lane < kFlowCopies && tau0 + p > kFlowLag) {
        uint64_t* tc =
            a.tot + (uint64_t((tau0 + p - kFlowLag) % kFlowTot) * kFlowCopies + lane) * kFlowStride;
        stAgent(tc, kCoopEmpty);
        stAgent(tc + 1, kCoopEmpty);
      }
    }
  }
}

// Reads the totals of iteration t - kFlowLag from copy `copy`: nu_{t-L} shifted into the history
// nuH (nuH[k] = nu_{t-L-k}), r_{t-2L} into rho. ra/rr: the words loaded already. false on abort.
__device__ inline bool flowTotals(const PiFlowArgs& a, const uint64_t* tc, uint64_t ra, uint64_t rr,
                                  double (&nuH)[kFlowLag + 1], double& rho) {
  for (uint32_t spin = 0; ra == kCoopEmpty || rr == kCoopEmpty; ++spin) {
    if (spin >= a.spinLimit || ldAgent(a.err)) return false;
    __builtin_amdgcn_s_sleep(1);
    ra = ldAgent(tc);
    rr = ldAgent(tc + 1);
  }
#pragma unroll
  for (uint32_t k = kFlowLag; k > 0; --k) nuH[k] = nuH[k - 1];
  nuH[0] = asF64(ra);
  rho = asF64(rr);
  return true;
}

__device__ inline bool flowStops(const PiFlowArgs& a, uint32_t t, double rho) {
  return t > 2 * kFlowLag && ((a.tol >= 0.0 && rho <= a.tol) || t - 2 * kFlowLag >= a.maxIter);
}

// FULL = false: every wave owns one group (the host checks) and no column has more than
// kCoopCache entries (every block checks *maxLen first and, if one has, all leave at once with
// *err = kFlowNeedsSweep, before any wait), so an iteration is straight-line code; FULL = true:
// waves sweep several groups and long columns read their further entries from the matrix every
// iteration.
template <bool FULL>
static __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void piFlowKernel(PiFlowArgs a) {
  const uint32_t b = blockIdx.x, lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if constexpr (!FULL) {
    if (*a.maxLen > uint32_t(kCoopCache)) {
      if (b == 0 && threadIdx.x == 0) stAgent(a.err, kFlowNeedsSweep);
      return;
    }
  }
  if (b == a.workers) {  // the aggregator: one wave
    if (w == 0) piFlowAggregator(a);
    return;
  }
  const uint32_t n = a.n, nG = a.nGroups, wave = 4 * b + w, waves = 4 * a.workers;
  if (wave >= nG) return;
  const uint32_t copy = wave % kFlowCopies;
  // the first group's column keeps its first kCoopCache entries for the whole run: row byte
  // offsets in registers, values in LDS (each lane reads only its own slots, so no barrier),
  // -0.0 past the column's end: the product of such a slot (its entry is never loaded: +0.0) is
  // -0.0, and x + (-0.0) == x for every x, so the slots past the end add nothing and no select
  // is needed; its own z history (z_{t-1}, ..., z_{t-1-kFlowLag}) stays in registers
  __shared__ double cvL[kCoopCache][kBlock];
  const uint32_t j0 = 64 * wave + lane;
  const bool valid0 = j0 < n;
  uint32_t c0 = 0, c1 = 0;
  if (valid0) {
    c0 = a.colPtr[j0];
    c1 = a.colPtr[j0 + 1];
  }
  uint32_t ci[kCoopCache];
#pragma unroll
  for (int u = 0; u < kCoopCache; ++u) {
    ci[u] = 0;
    cvL[u][threadIdx.x] = -0.0;
  }
  if (c1 > c0) {  // indices clamped into the column
#pragma unroll
    for (int u = 0; u < kCoopCache; ++u) {
      const uint32_t eu = min(c0 + u, c1 - 1);
      ci[u] = a.rowIdx[eu] * uint32_t(sizeof(double));  // n < 2^29
      if (c0 + u < c1) cvL[u][threadIdx.x] = a.val[eu];
    }
  }
  const uint32_t cnt0 = min(c1 - c0, uint32_t(kCoopCache));
  double own[kFlowLag + 1];
  const double z00 = valid0 ? a.mu0[j0] : 0.0;
#pragma unroll
  for (uint32_t k = 0; k <= kFlowLag; ++k) own[k] = z00;
  // nu_{t-kFlowLag-k} for k = 0 .. kFlowLag (nu_0 = 1 stands for the start)
  double nuH[kFlowLag + 1];
#pragma unroll
  for (uint32_t k = 0; k <= kFlowLag; ++k) nuH[k] = 1.0;
  const uint64_t* mu0b = reinterpret_cast<const uint64_t*>(a.mu0);
  if constexpr (!FULL) {
    // pending group sums of the previous iteration: |z_{t-1}| and its residual term's inputs
    double pAbs = 0.0, pZa = 0.0, pZb = 0.0, pNuA = 1.0, pNuB = 1.0;
    bool pRes = false;
    for (uint32_t t = 1;; ++t) {
      const uint64_t* src = t == 1 ? mu0b : a.z + uint64_t((t - 1) % kFlowRing) * n;
      // totals of t - L and the gathers of z_{t-1}: issued together
      const uint64_t* tc = a.tot + (uint64_t((t - kFlowLag) % kFlowTot) * kFlowCopies + copy) * kFlowStride;
      uint64_t ra = 0, rr = 0;
      if (t > kFlowLag) {
        ra = ldAgent(tc);
        rr = ldAgent(tc + 1);
      }
      uint64_t m0[kCoopCache];
      flowIssue<kCoopCache>(src, ci, cnt0, m0);
      // the previous iteration's group sums, while the loads are in flight
      if (t > 1) {
        const double term = (pRes && valid0) ? fabs(pZa / pNuA - pZb / pNuB) : 0.0;
        const double ga = waveButterflySum(pAbs);
        const double gr = waveButterflySum(term);
        if (lane == 0) {
          uint64_t* pc = a.part + (uint64_t((t - 1) % kFlowPart) * nG + wave) * 2;
          stAgent(pc, asU64(ga));
          stAgent(pc + 1, asU64(gr));
        }
      }
      if (t > kFlowLag) {
        double rho = 0.0;
        if (!flowTotals(a, tc, ra, rr, nuH, rho)) {
          if (lane == 0) flowAbort(a);
          return;
        }
        if (flowStops(a, t, rho)) return;
      }
      const double c = piScale(t, nuH[8 - kFlowLag]);  // nu_{t-8}
      if (!flowComplete<kCoopCache>(src, ci, cnt0, t > 1, m0, a)) {
        flowAbort(a);
        return;
      }
      const double s = flowCachedSum(m0, cvL, cnt0, c);
      __builtin_amdgcn_s_waitcnt(0);  // earlier stores (clears included) have completed
      if (valid0) {
        stAgent(a.z + uint64_t(t % kFlowRing) * n + j0, asU64(s));
        stAgent(a.z + uint64_t((t + kFlowRing - kFlowClear) % kFlowRing) * n + j0, kCoopEmpty);
      }
      // inputs of this iteration's group sums (formed in the next iteration): the residual
      // term of iteration t - L, |z_{t-L}/nu_{t-L} - z_{t-L-1}/nu_{t-L-1}|
      pAbs = valid0 ? fabs(s) : 0.0;
      pRes = t > kFlowLag;
      pZa = own[kFlowLag - 1];
      pZb = own[kFlowLag];
      pNuA = nuH[0];
      pNuB = nuH[1];
#pragma unroll
      for (uint32_t k = kFlowLag; k > 0; --k) own[k] = own[k - 1];
      own[0] = s;
    }
  } else {
    for (uint32_t t = 1;; ++t) {
      const uint64_t* src = t == 1 ? mu0b : a.z + uint64_t((t - 1) % kFlowRing) * n;
      uint64_t* dst = a.z + uint64_t(t % kFlowRing) * n;
      uint64_t* clr = a.z + uint64_t((t + kFlowRing - kFlowClear) % kFlowRing) * n;
      const bool check = t > 1;
      if (t > kFlowLag) {
        const uint64_t* tc = a.tot + (uint64_t((t - kFlowLag) % kFlowTot) * kFlowCopies + copy) * kFlowStride;
        double rho = 0.0;
        if (!flowTotals(a, tc, ldAgent(tc), ldAgent(tc + 1), nuH, rho)) {
          if (lane == 0) flowAbort(a);
          return;
        }
        if (flowStops(a, t, rho)) return;
      }
      const double c = piScale(t, nuH[8 - kFlowLag]);  // nu_{t-8}
      const double nuA = nuH[0], nuB = nuH[1];         // nu_{t-L}, nu_{t-L-1}
      for (uint32_t g = wave, first = 1; g < nG; g += waves, first = 0) {
        const uint32_t j = 64 * g + lane;
        const bool valid = j < n;
        double s = 0.0;
        uint32_t e, e1;
        if (first) {
          uint64_t m0[kCoopCache];
          flowIssue<kCoopCache>(src, ci, cnt0, m0);
          if (!flowComplete<kCoopCache>(src, ci, cnt0, check, m0, a)) {
            flowAbort(a);
            return;
          }
          s = flowCachedSum(m0, cvL, cnt0, 1.0);
          e = c0 + kCoopCache;
          e1 = c1;
        } else {
          e = valid ? a.colPtr[j] : 0;
          e1 = valid ? a.colPtr[j + 1] : 0;
        }
        // the rest of the column, kCoopUnroll entries at a time (sums selected)
        for (; __any(e < e1); e += kCoopUnroll) {
          const uint32_t cnt = e1 > e ? e1 - e : 0;
          uint32_t off[kCoopUnroll];
          double v[kCoopUnroll];
#pragma unroll
          for (int u = 0; u < kCoopUnroll; ++u) {
            const uint32_t eu = cnt ? min(e + u, e1 - 1) : 0;
            off[u] = cnt ? a.rowIdx[eu] * uint32_t(sizeof(double)) : 0;
            v[u] = cnt ? a.val[eu] : 0.0;
          }
          uint64_t m[kCoopUnroll];
          flowIssue<kCoopUnroll>(src, off, cnt, m);
          if (!flowComplete<kCoopUnroll>(src, off, cnt, check, m, a)) {
            flowAbort(a);
            return;
          }
#pragma unroll
          for (int u = 0; u < kCoopUnroll; ++u) {
            const double prod = asF64(m[u]) * v[u];
            s = (uint32_t(u) < cnt) ? s + prod : s;
          }
        }
        if (c != 1.0) s *= c;
        if (first) __builtin_amdgcn_s_waitcnt(0);  // earlier stores (clears included) have completed
        if (valid) {
          stAgent(dst + j, asU64(s));
          stAgent(clr + j, kCoopEmpty);
        }
        // the residual term of iteration t - L: |z_{t-L}/nu_{t-L} - z_{t-L-1}/nu_{t-L-1}|
        double za = 0.0, zb = 0.0;
        if (first) {
          za = own[kFlowLag - 1];
          zb = own[kFlowLag];
#pragma unroll
          for (uint32_t k = kFlowLag; k > 0; --k) own[k] = own[k - 1];
          own[0] = s;
        } else if (t > kFlowLag && valid) {
          za = asF64(ldAgent(a.z + uint64_t((t - kFlowLag) % kFlowRing) * n + j));
          zb = t - kFlowLag - 1 == 0 ? a.mu0[j]
                                     : asF64(ldAgent(a.z + uint64_t((t - kFlowLag - 1) % kFlowRing) * n + j));
        }
        const double term = (t > kFlowLag && valid) ? fabs(za / nuA - zb / nuB) : 0.0;
        const double ga = waveButterflySum(valid ? fabs(s) : 0.0);
        const double gr = waveButterflySum(term);
        if (lane == 0) {
          uint64_t* pc = a.part + (uint64_t(t % kFlowPart) * nG + g) * 2;
          stAgent(pc, asU64(ga));
          stAgent(pc + 1, asU64(gr));
        }
      }
    }
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

// Residual of the leading `lead` basis columns: per box sum over c < lead of (y_c - sum over
// d < lead of x_d H_dc)^2 (d, then c in order), summed over the boxes in dotPartialKernel's
// strided order; the block partials go to part[blockIdx.x]
template <int M>
static __global__ __launch_bounds__(kBlock) void residualPartialKernel(const double* X, const double* Y,
                                                                      const double* H, uint32_t n, int lead,
                                                                      double* part) {
  __shared__ double h[M * M];
  if (threadIdx.x < M * M) h[threadIdx.x] = H[threadIdx.x];
  __syncthreads();
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double x[M];
#pragma unroll
    for (int d = 0; d < M; ++d) x[d] = X[uint64_t(i) * M + d];
    double r = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (c < lead) {
        double v = Y[uint64_t(i) * M + c];
#pragma unroll
        for (int d = 0; d < M; ++d)
          if (d < lead) v -= x[d] * h[d * M + c];
        r += v * v;
      }
    }
    acc += r;
  }
  const double v = blockReduceSumF64(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// *out = the partials summed in sumPartials' order: one block
static __global__ __launch_bounds__(kBlock) void sumPartialsKernel(const double* part, uint32_t nb, double* out) {
  const double v = sumPartials(part, nb);
  if (threadIdx.x == 0) *out = v;
}

// H_ab = x_a . z_b from dotPartialKernel (SYM = 0): one block
static __global__ __launch_bounds__(kBlock) void ritzKernel(const double* part, uint32_t nb, int m, double* H) {
  for (int q = 0; q < m * m; ++q) {
    const double h = sumPartials(part + uint64_t(q) * nb, nb);
    if (threadIdx.x == 0) H[q] = h;
  }
}

// ---------------------------------------------------------------------------------------------
// Krylov basis for the Krylov-Schur eigensolver (build-defined: the reference leaves its
// eigenvalues to MATLAB eigs, subdivisionByNim.m:30-56). The basis is stored column by column
// (V[c * n + i]); the operator is the row-vector product w = v P (w_j = sum over column j in
// ascending row order, spmmKernel<1>). Every dot product over the boxes runs in the fixed order of
// dotPartialKernel (block-strided partials, blockReduceSumF64, sumPartials), so the results
// equal the oracle's restatement bit for bit.
// ---------------------------------------------------------------------------------------------
constexpr int kKrylovMaxCols = 64; // basis vectors at most (ncv + 1)

// The basis passes below hold one value per basis column in registers: NC is the column count
// rounded up to a multiple of 8 (8, 16, ..., 64). Lanes q >= cols load column cols - 1 again
// (lines the lane has just loaded) and their results are dropped by selects, so no branch
// depends on the column count. Every lane sums its rows i = t, t + T, ... (T = all lanes of
// the grid) in order, the block sums its lanes with blockReduceSumF64 and colDotsReduceKernel
// sums the block partials with sumPartials: the order of dotPartialKernel.

// block partials of x_c . w for every column c < cols (NC - 8 < cols <= NC) in one pass over V
// and w: part[c * gridDim.x + block]
template <int NC>
static __global__ __launch_bounds__(kBlock) void colDotsKernel(const double* __restrict__ V, uint32_t cols,
                                                              const double* __restrict__ w, uint32_t n,
                                                              double* __restrict__ part) {
  __builtin_assume(cols > uint32_t(NC - 8) && cols <= uint32_t(NC));
  double acc[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) acc[q] = 0.0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double x[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) x[q] = V[uint64_t(min(uint32_t(q), cols - 1)) * n + i];
    const double wi = w[i];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const double prod = x[q] * wi;
      const double a = acc[q] + prod;
      acc[q] = uint32_t(q) < cols ? a : acc[q];
    }
  }
#pragma unroll
  for (int q = 0; q < NC; ++q)
    if (uint32_t(q) < cols) {
      const double v = blockReduceSumF64(acc[q]);
      if (threadIdx.x == 0) part[uint64_t(q) * gridDim.x + blockIdx.x] = v;
    }
}

// Gram-Schmidt step and the next dot products in one pass over V: w_i = w_i - sum over c < cols
// of x_c[i] h_c (c ascending, one subtraction per column), stored; then, with NORM = false, the
// block partials of x_c . w for c < cols (part[c * gridDim.x + block]), or, with NORM = true, of
// w . w (part[block]), both of the updated w.
template <int NC, bool NORM>
static __global__ __launch_bounds__(kBlock) void colUpdateKernel(const double* __restrict__ V, uint32_t cols,
                                                                const double* __restrict__ h, double* __restrict__ w,
                                                                uint32_t n, double* __restrict__ part) {
  __builtin_assume(cols > uint32_t(NC - 8) && cols <= uint32_t(NC));
  constexpr int NA = NORM ? 1 : NC;
  double acc[NA];
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = 0.0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double x[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) x[q] = V[uint64_t(min(uint32_t(q), cols - 1)) * n + i];
    double v = w[i];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const double u = v - x[q] * h[min(uint32_t(q), cols - 1)];
      v = uint32_t(q) < cols ? u : v;
    }
    w[i] = v;
    if (NORM) {
      const double prod = v * v;
      acc[0] += prod;
    } else {
#pragma unroll
      for (int q = 0; q < NA; ++q) {
        const double prod = x[q] * v;
        const double a = acc[q] + prod;
        acc[q] = uint32_t(q) < cols ? a : acc[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NA; ++q)
    if (NORM || uint32_t(q) < cols) {
      const double v = blockReduceSumF64(acc[q]);
      if (threadIdx.x == 0) part[uint64_t(q) * gridDim.x + blockIdx.x] = v;
    }
}

// h[c] = the partials of column c summed in sumPartials' order: one block per column
static __global__ __launch_bounds__(kBlock) void colDotsReduceKernel(const double* part, uint32_t nb, double* h) {
  const double v = sumPartials(part + uint64_t(blockIdx.x) * nb, nb);
  if (threadIdx.x == 0) h[blockIdx.x] = v;
}

// out_i = w_i / *beta
static __global__ __launch_bounds__(kBlock) void scaleIntoKernel(const double* w, uint32_t n, const double* beta,
                                                                double* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = w[i] / *beta;
}

// V[:, c] = sum over d < m of V[:, d] Q[d * k + c] (d ascending, from 0.0) for c < k, in place
// (NC - 8 < k <= NC). Q (row-major m x k) is staged in LDS as rows of NC with zeros past k; a
// lane streams its box's m entries into NC sums (those past k are dropped) and writes the k
// sums after its last read of the box's row. Lanes stride over the boxes, so at most
// kRotateBlocks blocks stage Q.
constexpr uint32_t kRotateBlocks = 4096;

template <int NC>
static __global__ __launch_bounds__(kBlock) void krylovRotateKernel(double* V, uint32_t n, int m, int k,
                                                                   const double* __restrict__ Q) {
  __shared__ double q[kKrylovMaxCols * NC];
  for (int t = threadIdx.x; t < m * NC; t += kBlock) {
    const int d = t / NC, c = t % NC;
    q[t] = c < k ? Q[d * k + c] : 0.0;
  }
  __syncthreads();
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = 0.0;
    for (int d = 0; d < m; ++d) {
      const double x = V[uint64_t(d) * n + i];
#pragma unroll
      for (int c = 0; c < NC; ++c) s[c] += x * q[d * NC + c];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c < k) V[uint64_t(c) * n + i] = s[c];
  }
}

// the default start vector: v_i = (splitmix64(8 i + 1) >> 11) 2^-53 * 2 - 1 (the subspace
// iteration's second start column)
static __global__ __launch_bounds__(kBlock) void krylovStartKernel(double* v, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = double(splitmix64(8 * uint64_t(i) + 1) >> 11) * 0x1p-53 * 2.0 - 1.0;
}

// box-by-box basis to column-by-column (the C ABI's layout)
static __global__ __launch_bounds__(kBlock) void transposeBasisKernel(const double* X, uint32_t n, int m,
                                                                     double* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < m; ++c) out[uint64_t(c) * n + i] = X[uint64_t(i) * m + c];
}

// Eight independent FP64 multiply-add chains per lane (no memory traffic in the loop); each
// block's first lane records the shader-clock and 100 MHz real-time counters around its loop, so
// the host can read the clock the device runs at under a full FP64 load.
static __global__ __launch_bounds__(kBlock) void fp64ClockKernel(double* out, uint64_t* clk, int iters, double s) {
  double a0 = threadIdx.x, a1 = a0 + 1.0, a2 = a0 + 2.0, a3 = a0 + 3.0, a4 = a0 + 4.0, a5 = a0 + 5.0,
         a6 = a0 + 6.0, a7 = a0 + 7.0;
  const uint64_t t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < iters; ++i) {
    a0 = a0 * s + 1.0; a1 = a1 * s + 1.0; a2 = a2 * s + 1.0; a3 = a3 * s + 1.0;
    a4 = a4 * s + 1.0; a5 = a5 * s + 1.0; a6 = a6 * s + 1.0; a7 = a7 * s + 1.0;
  }
  const uint64_t t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  out[uint64_t(blockIdx.x) * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
  if (threadIdx.x == 0) {
    clk[2 * blockIdx.x] = t1 - t0;
    clk[2 * blockIdx.x + 1] = r1 - r0;
  }
}

// Plain copy (traffic-counter calibration): every byte read once and written once.
static __global__ __launch_bounds__(kBlock) void copyKernel(const uint4* in, uint4* out, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock)
    out[i] = in[i];
}

// *out = max(colPtr[j + 1] - colPtr[j]) over the columns (out zero at launch); one atomic per block
static __global__ __launch_bounds__(kBlock) void colMaxLenKernel(const uint32_t* colPtr, uint32_t n, uint32_t* out) {
  __shared__ uint32_t waveMax[kBlock / kWave];
  uint32_t m = 0;
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) m = max(m, colPtr[j + 1] - colPtr[j]);
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) waveMax[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 1; w < kBlock / kWave; ++w) m = max(m, waveMax[w]);
    if (m) atomicMax(out, m);
  }
}

// The start of a solve in one launch: the state, mu_0 = 1/n unless the caller gave it, and for
// the persistent form (ring != nullptr) its ring, partials and totals at the empty pattern and
// its error flag and longest-column slot at zero.
struct PiInitArgs {
  PiState* st;
  PiState init;
  double* mu0;
  uint32_t n;
  int fillMu;
  uint64_t* ring;
  uint64_t nRing;
  uint64_t* part;
  uint64_t nPart;
  uint64_t* tot;
  int* err;
  uint32_t* maxLen;
};

static __global__ __launch_bounds__(kBlock) void piInitKernel(PiInitArgs a) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  const uint64_t first = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (first == 0) {
    *a.st = a.init;
    if (a.ring) {
      *a.err = 0;
      *a.maxLen = 0;
    }
  }
  if (a.fillMu) {
    const double v = 1.0 / double(a.n);
    for (uint64_t i = first; i < a.n; i += stride) a.mu0[i] = v;
  }
  if (!a.ring) return;
  for (uint64_t i = first; i < a.nRing; i += stride) a.ring[i] = kCoopEmpty;
  for (uint64_t i = first; i < a.nPart; i += stride) a.part[i] = kCoopEmpty;
  for (uint64_t i = first; i < kFlowTotWords; i += stride) a.tot[i] = kCoopEmpty;
}

static __global__ __launch_bounds__(kBlock) void piFinishKernel(const double* z, const double* mu0, uint32_t n,
                                                         const PiState* st, double* mu) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  mu[j] = st->final == 0 ? mu0[j] : z[uint64_t(st->slot) * n + j] / st->nuFinal;
}

}  // namespace b12x
