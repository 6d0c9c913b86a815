// Kernels of the forest: many jobs' trees in one set of leaf arrays, so that one launch per
// phase serves every job (batched parameter sweeps and continuous-batching serving).
//
// Layout: the leaves of a job are contiguous and sorted by key; job[i] names the job slot of
// leaf i. Per-slot tables hold the job's domain and map parameters (JobDev) and its leaf range
// [off, end), recomputed on the device after every compaction by forestBoundaryKernel. A test
// point descends only within its own job's range, and grid indices are job-local, so a job's
// result does not depend on which other jobs share the launch.
#pragma once

#include "b12x_kernels.hpp"

namespace b12x {

struct JobDev {
  Domain dom;
  MapParams map;
};

struct ForestView {
  uint64_t* key;
  uint8_t* depth;
  uint8_t* flags;
  uint32_t* job;
};

// subdivide(SD) of every leaf of every job (ImplicitBoxTree.hpp:2883-2930); children inherit the
// job id.
struct ForestSubdivideOp : OpBase {
  ForestView in, out;
  uint8_t mask;
  __device__ uint32_t count(uint32_t i) const {
    return (areFlagsSet(in.flags[i], mask) && in.depth[i] < kMaxDepth) ? 2u : 1u;
  }
  __device__ void emit(uint32_t i, uint32_t c, uint32_t pos) const {
    const uint64_t k = in.key[i];
    const int d = in.depth[i];
    const uint8_t f = in.flags[i];
    const uint32_t j = in.job[i];
    if (c == 1) {
      out.key[pos] = k;
      out.depth[pos] = uint8_t(d);
      out.flags[pos] = f;
      out.job[pos] = j;
    } else {
      const uint64_t b = bitsOf(k, d) << 1;
      out.key[pos] = keyOf(b, d + 1);
      out.key[pos + 1] = keyOf(b | 1ull, d + 1);
      out.depth[pos] = out.depth[pos + 1] = uint8_t(d + 1);
      out.flags[pos] = out.flags[pos + 1] = f;
      out.job[pos] = out.job[pos + 1] = j;
    }
  }
};

// mode 0: remove(HIT) + unsetFlags(HIT) of the subdivision step (hit bytes mark HIT);
// mode 1: drop the leaves of retiring jobs (retiring[job] != 0).
struct ForestKeepOp : OpBase {
  ForestView in, out;
  uint8_t* hit;
  const uint8_t* retiring;
  int mode;
  __device__ uint32_t count(uint32_t i) const {
    if (mode == 1) return retiring[in.job[i]] ? 0u : 1u;
    const uint8_t f = in.flags[i] | (hit[i] ? uint8_t(F_HIT) : uint8_t(0));
    return areFlagsSet(f, F_HIT) ? 1u : 0u;
  }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const {
    out.key[pos] = in.key[i];
    out.depth[pos] = in.depth[i];
    out.flags[pos] = mode == 1 ? in.flags[i] : uint8_t(in.flags[i] & ~F_HIT);
    out.job[pos] = in.job[i];
  }
  __device__ void after(uint32_t i) const {
    if (mode == 0) hit[i] = 0;
  }
};

// indices of the leaves of retiring jobs (source list of the batched transition matrix)
struct ForestListOp : OpBase {
  const uint32_t* job;
  const uint8_t* retiring;
  uint32_t* out;
  __device__ uint32_t count(uint32_t i) const { return retiring[job[i]] ? 1u : 0u; }
  __device__ void emit(uint32_t i, uint32_t, uint32_t pos) const { out[pos] = i; }
};

static __global__ __launch_bounds__(kBlock) void forestClearRangesKernel(uint32_t* off, uint32_t* end, uint32_t nSlots) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j < nSlots) off[j] = end[j] = 0;
}

// [off, end) of every job present in the arrays (jobs are contiguous)
static __global__ __launch_bounds__(kBlock) void forestBoundaryKernel(const uint32_t* job, uint32_t n, const uint32_t* nDev,
                                                               uint32_t* off, uint32_t* end) {
  n = nDev ? *nDev : n;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = job[i];
  if (i == 0 || job[i - 1] != j) off[j] = i;
  if (i == n - 1 || job[i + 1] != j) end[j] = i + 1;
}

struct ForestImageArgs {
  const JobDev* jobs;
  const uint32_t* off;
  const uint32_t* end;
  const uint64_t* key;
  const uint8_t* depth;
  const uint32_t* job;
  const uint32_t* srcList;  // optional: source leaves (else every leaf)
  uint32_t nSrc;            // number of source leaves (upper bound when nSrcDev is set)
  const uint32_t* nSrcDev;
  GridParams grid;          // kind and k shared by the forest (InnerGrid table shared too)
  int mode;                 // 0: hit bytes, 1: out[i] = global target leaf or kInvalid
  uint8_t* hit;
  uint32_t* out;
};

template <int DIM, int MAP, int GRID>
static __global__ __launch_bounds__(kBlock) void forestImageKernel(ForestImageArgs a) {
  const uint32_t nSrc = a.nSrcDev ? *a.nSrcDev : a.nSrc;
  const uint64_t nPoints = uint64_t(nSrc) * a.grid.k;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < nPoints; i += stride) {
    const uint32_t b = uint32_t(i / a.grid.k);
    const uint32_t g = uint32_t(i - uint64_t(b) * a.grid.k);
    const uint32_t leaf = a.srcList ? a.srcList[b] : b;
    const uint32_t j = a.job[leaf];
    const JobDev& J = a.jobs[j];
    const uint32_t o = a.off[j], e = a.end[j];
    double c[DIM], rb[DIM], gp[DIM], x[DIM], y[DIM];
    boxCoords<DIM>(J.dom, a.key[leaf], a.depth[leaf], c, rb);
    if constexpr (GRID == GRID_INNER) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) gp[d] = a.grid.table[g * DIM + d];
    } else {
      uniformGridPoint<DIM>(a.grid.seed, uint64_t(leaf - o) * a.grid.k + g, gp);  // job-local index
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = c[d] + gp[d] * rb[d];
    applyMap<DIM, MAP>(J.map, x, y);
    const uint64_t pk = pointKey<DIM>(J.dom, y);
    const uint32_t t = locate(a.key + o, a.depth + o, e - o, pk);
    if (a.mode == 0) {
      if (t != kInvalid) a.hit[o + t] = 1;
    } else {
      a.out[i] = t == kInvalid ? kInvalid : o + t;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Batched transition matrices: entries of all retiring jobs, rows = positions in the source
// list (job-major), columns = global leaf indices. Local indices per job are row - lo[r] and
// col - off[slot].
// ---------------------------------------------------------------------------------------------
static __global__ __launch_bounds__(kBlock) void forestLocalizeKernel(const uint32_t* rowPos, const uint32_t* col,
                                                               uint64_t nnz, const uint32_t* jobOfPos,
                                                               const uint32_t* loOfR, const uint32_t* offOfR,
                                                               uint32_t* localRow, uint32_t* localCol,
                                                               uint32_t* colPos) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nnz) return;
  const uint32_t p = rowPos[e];
  const uint32_t r = jobOfPos[p];
  const uint32_t lc = col[e] - offOfR[r];
  localRow[e] = p - loOfR[r];
  localCol[e] = lc;
  colPos[e] = loOfR[r] + lc;  // column as a list position (same job)
}

// ---------------------------------------------------------------------------------------------
// Power-iteration pool: the transition matrices of finished jobs, block-diagonal over pool
// positions (a job's rows and columns are its own contiguous position range [lo, hi)). Jobs join
// and leave at any time and each carries its own iteration counter t, norm ring, residual and
// stopping decision. A job's sums run over its own range in a fixed order, so its measure does
// not depend on the other jobs in the pool.
// ---------------------------------------------------------------------------------------------
// Blocks of poolStepKernel cover kPoolCols consecutive columns of one job (job-relative chunks),
// so each block's partial sums depend only on that job's data.
constexpr uint32_t kPoolCols = kBlock;
// Pool entries store the hit count (1 byte); the value count / k comes from a table of the
// IEEE quotients double(c) / double(k), c = 0..64 (forests take k <= 64).
constexpr uint32_t kPoolTable = 65;

struct PoolJobDev {
  uint32_t lo, hi;  // pool positions
  uint32_t t;       // next iteration of this job
  int32_t done;
  uint32_t final;
  uint32_t b0;      // first step-kernel block of this job
  double norm[3];
  double resid;
  double nuLag;     // nu_t of the last t with t % 8 == 0 (the scale c_{t+8} derives from it)
  uint32_t arrived; // blocks of the current batched step that have stored their partials (0 at rest)
};

__host__ __device__ inline uint32_t poolBlocksOf(uint32_t len) { return (len + kPoolCols - 1) / kPoolCols; }

// jobOfBlock[b] = r for the blocks of pool jobs r in [r0, r1)
static __global__ __launch_bounds__(kBlock) void poolJobOfBlockKernel(const PoolJobDev* jobs, uint32_t r0, uint32_t r1,
                                                                     uint32_t* jobOfBlock) {
  const uint32_t r = r0 + blockIdx.x;
  if (r >= r1) return;
  const uint32_t b0 = jobs[r].b0, nb = poolBlocksOf(jobs[r].hi - jobs[r].lo);
  for (uint32_t q = threadIdx.x; q < nb; q += kBlock) jobOfBlock[b0 + q] = r;
}

// Canonical sum of a job's block partials q = 0..nb-1 (shared by the batched and the persistent
// forms, so a job's iterates do not depend on which form ran them): lane l of one wave adds
// q = l, l + 64, ... in ascending order, then the lanes combine by a fixed shuffle tree.
__device__ inline double poolLaneTreeSum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// One wave (the lanes 0..63 of the calling block) closes iteration t of job J: sums its block
// partials (agent-scope loads) in the canonical order (norm of y_t, residual of iteration t-1), stopping decision,
// t += 1; the arrival count returns to 0.
__device__ inline void poolReduceJob(const double* partAbs, const double* partRes, PoolJobDev& J, uint32_t t,
                                     uint32_t maxIter, double tol) {
  const uint32_t b0 = J.b0, nb = poolBlocksOf(J.hi - J.lo), lane = threadIdx.x & (kWave - 1);
  constexpr int U = 8;  // loads in flight per lane before they are added (order unchanged)
  double ta = 0.0, tr = 0.0;
  for (uint32_t q0 = 0; q0 < nb; q0 += kWave * U) {
    double va[U], vr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t q = q0 + u * kWave + lane;
      va[u] = q < nb ? ldAgent(partAbs + b0 + q) : 0.0;
      vr[u] = q < nb ? ldAgent(partRes + b0 + q) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (q0 + u * kWave + lane < nb) {
        ta += va[u];
        tr += vr[u];
      }
  }
  ta = poolLaneTreeSum(ta);
  tr = poolLaneTreeSum(tr);
  if (lane == 0) {
    J.norm[t % 3] = ta;
    if ((t & 7u) == 0) J.nuLag = ta;
    const uint32_t completed = t - 1;
    if (completed >= 1) {
      J.resid = tr;
      if ((tol >= 0.0 && tr <= tol) || completed >= maxIter) {
        J.done = 1;
        J.final = completed;
      }
    }
    J.t = t + 1;
    J.arrived = 0;
  }
}

// y_t for the kPoolCols columns of pool block v (job J, iteration t), one lane per column: the
// power iteration of b12x_kernels.hpp (unnormalised iterates y_t = c_t y_{t-1} P, c_t = piScale)
// over the job's positions. The lane adds its column's products in ascending-row order (the
// oracle's order) with kPoolUnroll entries' loads in flight at a time. Returns the lane's terms
// of the block's partial sums: |y_t| and the residual term of iteration t-1,
// |y_{t-1} / nu_{t-1} - y_{t-2} / nu_{t-2}| (y_0 = mu_0 uniform, nu_0 = 1).
constexpr int kPoolUnroll = 8;

__device__ inline void poolColumnsStep(const uint32_t* __restrict__ colPtr, const uint32_t* __restrict__ rowIdx,
                                       const uint8_t* __restrict__ cnt, const double* val, const PoolJobDev& J,
                                       uint32_t v, uint32_t t, double nPrev, double nPrev2, double nuLag, double* y,
                                       uint32_t ys, double& sAbs, double& sRes) {
  const uint32_t s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;
  const double mu0 = 1.0 / double(J.hi - J.lo);
  const double c = piScale(t, nuLag);
  const uint32_t p = J.lo + (v - J.b0) * kPoolCols + threadIdx.x;
  const double* yPrev = y + uint64_t(s1) * ys;
  const double* yPrev2 = y + uint64_t(s2) * ys;
  sAbs = 0.0;
  sRes = 0.0;
  if (p >= J.hi) return;
  const uint32_t e1 = colPtr[p + 1];
  double s = 0.0;
  // branch-free body (entry indices clamped to the column, values and sums selected) so that
  // all of a lane's gathers are in flight at once; y is always allocated, so the unused read of
  // iteration 1 stays in bounds
  for (uint32_t e = colPtr[p]; e < e1; e += kPoolUnroll) {
    uint32_t i[kPoolUnroll];
    double w[kPoolUnroll], m[kPoolUnroll];
#pragma unroll
    for (int u = 0; u < kPoolUnroll; ++u) {
      const uint32_t eu = min(e + u, e1 - 1);
      i[u] = rowIdx[eu];
      w[u] = val[cnt[eu]];
    }
#pragma unroll
    for (int u = 0; u < kPoolUnroll; ++u) m[u] = yPrev[i[u]];
#pragma unroll
    for (int u = 0; u < kPoolUnroll; ++u) {
      const double z = (t == 1) ? mu0 : m[u];
      const double prod = z * w[u];
      s = (e + u < e1) ? s + prod : s;
    }
  }
  if (c != 1.0) s *= c;
  y[uint64_t(s0) * ys + p] = s;
  sAbs = fabs(s);
  if (t >= 2) {
    const double x = yPrev[p] / nPrev;
    const double z = (t == 2) ? mu0 : yPrev2[p] / nPrev2;
    sRes = fabs(x - z);
  }
}

// Iteration t of every pool job in one launch, one block per pool block (poolColumnsStep). The
// block stores its partial sums with agent-scope stores, waits for them to complete and counts
// itself in with an agent-scope atomic; the job's last block to arrive closes the iteration with
// poolReduceJob, whose agent-scope loads see every block's partials. No block waits for another.
// (Acquire-release fences at agent scope would write back and invalidate the XCD's L2 per block:
// four times slower measured.)
static __global__ __launch_bounds__(kBlock) void poolStepKernel(const uint32_t* __restrict__ colPtr,
                                                               const uint32_t* __restrict__ rowIdx,
                                                               const uint8_t* __restrict__ cnt,
                                                               const double* __restrict__ table,
                                                               const uint32_t* __restrict__ jobOfBlock,
                                                               PoolJobDev* jobs, double* y, uint32_t ys,
                                                               double* partAbs, double* partRes, uint32_t maxIter,
                                                               double tol) {
  __shared__ double val[kPoolTable];
  __shared__ int closes;
  PoolJobDev& J = jobs[jobOfBlock[blockIdx.x]];
  if (J.done) return;  // uniform over the block
  if (threadIdx.x < kPoolTable) val[threadIdx.x] = table[threadIdx.x];
  __syncthreads();
  const uint32_t t = J.t;
  const double nPrev = t >= 2 ? J.norm[(t + 2) % 3] : 1.0;
  const double nPrev2 = t >= 3 ? J.norm[(t + 1) % 3] : 1.0;
  const uint32_t nb = poolBlocksOf(J.hi - J.lo);
  double sAbs, sRes;
  poolColumnsStep(colPtr, rowIdx, cnt, val, J, blockIdx.x, t, nPrev, nPrev2, J.nuLag, y, ys, sAbs, sRes);
  sAbs = blockReduceSumF64(sAbs);
  sRes = blockReduceSumF64(sRes);
  if (threadIdx.x == 0) {
    stAgent(partAbs + blockIdx.x, sAbs);
    stAgent(partRes + blockIdx.x, sRes);
    __builtin_amdgcn_s_waitcnt(0);  // the partials have completed before the block counts itself in
    closes = __hip_atomic_fetch_add(&J.arrived, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nb - 1;
  }
  __syncthreads();
  if (closes && threadIdx.x < kWave) poolReduceJob(partAbs, partRes, J, t, maxIter, tol);
}

// Persistent form of the pool for pools of at most kPoolCoopMaxBlocks step blocks (every block
// resident): each job's blocks iterate together, synchronised only with one another through
// the job's partial slots (buffer t % 3, kCoopEmpty until written, cleared one iteration ahead,
// as the persistent forms of b12x_kernels.hpp use them), so jobs advance independently. Per
// launch a job runs until it stops or `budget` iterations are done; its first block then stores
// the job state. A wait past spinLimit polls raises *err and every block leaves at its next
// wait; the blocks of one job can then stand one iteration apart, so the host restarts the jobs
// still running from iteration 1 in the batched form (their iterates do not depend on the form).
// Same arithmetic and summation orders as poolStepKernel (and its poolReduceJob).
constexpr uint32_t kPoolCoopMaxBlocks = 512;
constexpr int kPoolCache = 24;  // entries of a lane's column held in registers

struct PoolCoopArgs {
  const uint32_t* colPtr;
  const uint32_t* rowIdx;
  const uint8_t* cnt;
  const double* table;  // kPoolTable values
  const uint32_t* jobOfBlock;
  PoolJobDev* jobs;
  double* y;
  uint32_t ys;
  uint64_t* part;  // 3 buffers x nBlocks slots x 2, kCoopEmpty at launch
  uint32_t nBlocks;
  int* err;        // zero at launch
  uint32_t maxIter;
  double tol;
  uint32_t budget;
  uint32_t spinLimit;
};

static __global__ __launch_bounds__(kBlock) void poolCoopKernel(PoolCoopArgs a) {
  constexpr uint32_t kSlots = kPoolCoopMaxBlocks / kWave;
  const uint32_t b = blockIdx.x, tid = threadIdx.x, G = a.nBlocks;
  __shared__ double sums[2];
  __shared__ int leave;
  const uint32_t r = a.jobOfBlock[b];
  PoolJobDev* J = a.jobs + r;
  if (J->done) return;
  const uint32_t lo = J->lo, hi = J->hi, b0 = J->b0, nb = poolBlocksOf(hi - lo);
  uint32_t t = J->t;
  double nPrev = t >= 2 ? J->norm[(t + 2) % 3] : 1.0;
  double nPrev2 = t >= 3 ? J->norm[(t + 1) % 3] : 1.0;
  double nuLag = J->nuLag;
  const double mu0 = 1.0 / double(hi - lo);
  const uint32_t p = lo + (b - b0) * kPoolCols + tid;
  const bool valid = p < hi;
  // the lane's column: up to kPoolCache entries in registers, the rest read per iteration
  uint32_t c0 = 0, c1 = 0;
  uint32_t ci[kPoolCache];
  double cv[kPoolCache];
#pragma unroll
  for (int u = 0; u < kPoolCache; ++u) {
    ci[u] = 0;
    cv[u] = 0.0;
  }
  if (valid) {
    c0 = a.colPtr[p];
    c1 = a.colPtr[p + 1];
    if (c1 > c0) {
#pragma unroll
      for (int u = 0; u < kPoolCache; ++u) {
        const uint32_t eu = min(c0 + u, c1 - 1);
        ci[u] = a.rowIdx[eu];
        cv[u] = a.table[a.cnt[eu]];
      }
    }
  }
  for (uint32_t m = 0;; ++m) {
    if (tid == 0) {
      uint64_t* nextSlot = a.part + (uint64_t((t + 1) % 3) * G + b) * 2;
      stAgent(nextSlot, kCoopEmpty);
      stAgent(nextSlot + 1, kCoopEmpty);
    }
    const double* yPrev = a.y + uint64_t((t + 2) % 3) * a.ys;
    const double* yPrev2 = a.y + uint64_t((t + 1) % 3) * a.ys;
    double* yCur = a.y + uint64_t(t % 3) * a.ys;
    double sAbs = 0.0, sRes = 0.0;
    const double c = piScale(t, nuLag);
    if (valid) {
      const double yp = ldAgent(yPrev + p), yp2 = ldAgent(yPrev2 + p);
      double mv[kPoolCache];
#pragma unroll
      for (int u = 0; u < kPoolCache; ++u) mv[u] = 0.0;
#pragma unroll
      for (int g = 0; g < kPoolCache; g += 8) {
        if (__any(c0 + g < c1)) {
#pragma unroll
          for (int u = g; u < g + 8; ++u) mv[u] = ldAgent(yPrev + ci[u]);
        }
      }
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < kPoolCache; ++u) {
        const double z = (t == 1) ? mu0 : mv[u];
        const double prod = z * cv[u];
        s = (c0 + u < c1) ? s + prod : s;
      }
      for (uint32_t e = c0 + kPoolCache; e < c1; e += 8) {
        uint32_t i[8];
        double v[8], mm[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint32_t eu = min(e + u, c1 - 1);
          i[u] = a.rowIdx[eu];
          v[u] = a.table[a.cnt[eu]];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) mm[u] = ldAgent(yPrev + i[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double z = (t == 1) ? mu0 : mm[u];
          const double prod = z * v[u];
          s = (e + u < c1) ? s + prod : s;
        }
      }
      if (c != 1.0) s *= c;
      stAgent(yCur + p, s);
      sAbs = fabs(s);
      if (t >= 2) {
        const double x = yp / nPrev;
        const double z = (t == 2) ? mu0 : yp2 / nPrev2;
        sRes = fabs(x - z);
      }
    }
    sAbs = blockReduceSumF64(sAbs);
    sRes = blockReduceSumF64(sRes);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's y stores have completed
    __syncthreads();
    uint64_t* buf = a.part + uint64_t(t % 3) * G * 2;
    if (tid == 0) {
      stAgent(buf + 2 * b, uint64_t(__double_as_longlong(sAbs)));
      stAgent(buf + 2 * b + 1, uint64_t(__double_as_longlong(sRes)));
    }
    if (tid < kWave) {
      uint64_t va[kSlots], vr[kSlots];
#pragma unroll
      for (uint32_t k = 0; k < kSlots; ++k) va[k] = vr[k] = 0;
      bool ok = false;
      for (uint32_t spin = 0; spin < a.spinLimit; ++spin) {
        bool mine = true;
#pragma unroll
        for (uint32_t k = 0; k < kSlots; ++k) {
          const uint32_t q = tid + k * kWave;
          if (q < nb && (va[k] == kCoopEmpty || vr[k] == kCoopEmpty || spin == 0)) {
            va[k] = ldAgent(buf + 2 * (b0 + q));
            vr[k] = ldAgent(buf + 2 * (b0 + q) + 1);
          }
          mine = mine && (q >= nb || (va[k] != kCoopEmpty && vr[k] != kCoopEmpty));
        }
        ok = __all(mine);
        if (ok || ldAgent(a.err)) break;
        __builtin_amdgcn_s_sleep(1);
      }
      double ta = 0.0, tr = 0.0;
#pragma unroll
      for (uint32_t k = 0; k < kSlots; ++k)
        if (tid + k * kWave < nb) {
          ta += __longlong_as_double((long long)va[k]);
          tr += __longlong_as_double((long long)vr[k]);
        }
      ta = poolLaneTreeSum(ta);
      tr = poolLaneTreeSum(tr);
      if (tid == 0) {
        if (!ok) stAgent(a.err, 1);
        leave = ok ? 0 : 1;
        sums[0] = ta;
        sums[1] = tr;
      }
    }
    __syncthreads();
    if (leave) return;  // the host restarts the job
    const double ta = sums[0], tr = sums[1];
    if ((t & 7u) == 0) nuLag = ta;
    const uint32_t completed = t - 1;
    const bool done = completed >= 1 && ((a.tol >= 0.0 && tr <= a.tol) || completed >= a.maxIter);
    if (done || m + 1 >= a.budget) {
      if (b == b0 && tid == 0) {
        J->norm[t % 3] = ta;
        J->norm[(t + 2) % 3] = nPrev;
        J->norm[(t + 1) % 3] = nPrev2;
        J->nuLag = nuLag;
        if (completed >= 1) J->resid = tr;
        if (done) {
          J->done = 1;
          J->final = completed;
        }
        J->t = t + 1;
      }
      return;
    }
    nPrev2 = nPrev;
    nPrev = ta;
    ++t;
    __syncthreads();  // sums[] is rewritten by the next wait
  }
}

// mu of a finished pool job into out[0 .. hi-lo)
static __global__ __launch_bounds__(kBlock) void poolFinishKernel(const double* y, uint32_t ys, const PoolJobDev* jobs,
                                                                 uint32_t r, double* out) {
  const PoolJobDev& J = jobs[r];
  const uint32_t len = J.hi - J.lo;
  const uint32_t f = J.final;
  for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < len; q += gridDim.x * kBlock) {
    const uint32_t p = J.lo + q;
    out[q] = f == 0 ? 1.0 / double(len) : y[uint64_t(f % 3) * ys + p] / J.norm[f % 3];
  }
}

// Pool compaction: position p of a kept job moves to p - shiftOfJob[job]; entries likewise.
static __global__ __launch_bounds__(kBlock) void poolMoveColumnsKernel(const uint32_t* colPtrIn, const uint32_t* jobOfPosIn,
                                                                      const double* yIn, uint32_t nIn, uint32_t ysIn,
                                                                      const uint32_t* posShift, const uint32_t* nnzShift,
                                                                      const int* keep, const uint32_t* newIndex,
                                                                      uint32_t* colPtrOut, uint32_t* jobOfPosOut,
                                                                      double* yOut, uint32_t ysOut) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nIn) return;
  const uint32_t r = jobOfPosIn[p];
  if (!keep[r]) return;
  const uint32_t q = p - posShift[r];
  colPtrOut[q] = colPtrIn[p] - nnzShift[r];
  jobOfPosOut[q] = newIndex[r];
  for (int sl = 0; sl < 3; ++sl) yOut[uint64_t(sl) * ysOut + q] = yIn[uint64_t(sl) * ysIn + p];
}

static __global__ __launch_bounds__(kBlock) void poolMoveEntriesKernel(const uint32_t* rowIn, const uint8_t* valIn,
                                                                      uint64_t nnzIn, const uint32_t* jobOfPosIn,
                                                                      const uint32_t* posShift, const uint32_t* nnzShift,
                                                                      const int* keep, uint32_t* rowOut,
                                                                      uint8_t* valOut) {
  const uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nnzIn) return;
  const uint32_t row = rowIn[e];
  const uint32_t r = jobOfPosIn[row];  // rows and columns of an entry belong to the same job
  if (!keep[r]) return;
  rowOut[e - nnzShift[r]] = row - posShift[r];
  valOut[e - nnzShift[r]] = valIn[e];
}

static __global__ __launch_bounds__(kBlock) void jobOfPosKernel(const uint32_t* lo, uint32_t nR, uint32_t* jobOfPos) {
  const uint32_t r = blockIdx.x;
  if (r >= nR) return;
  for (uint32_t p = lo[r] + threadIdx.x; p < lo[r + 1]; p += kBlock) jobOfPos[p] = r;
}

}  // namespace b12x
