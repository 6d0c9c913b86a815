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
// Block-diagonal power iteration over the retiring jobs (columns and rows are list positions).
// Each job has its own norm, residual and stopping decision; a job's sums are taken over its own
// column range in a fixed order, so its measure does not depend on the other jobs.
// ---------------------------------------------------------------------------------------------
struct BatchPiState {
  uint32_t t;      // iteration of the next step launch
  uint32_t tR;     // iteration being reduced
  uint32_t nDone;
  uint32_t pad;
};

static __global__ __launch_bounds__(kBlock) void batchPiStepKernel(const uint32_t* __restrict__ colPtr,
                                                            const uint32_t* __restrict__ rowIdx,
                                                            const double* __restrict__ val,
                                                            const uint32_t* __restrict__ jobOfPos,
                                                            const uint32_t* __restrict__ lo, uint32_t nR,
                                                            double* y, uint32_t n, const double* norm /*[3][nR]*/,
                                                            const int* done, double* absCol, double* resCol,
                                                            BatchPiState* st) {
  if (st->nDone >= nR) return;
  const uint32_t t = st->t;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->tR = t;
  const uint32_t s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;
  const double* yPrev = y + uint64_t(s1) * n;
  const double* yPrev2 = y + uint64_t(s2) * n;
  double* yCur = y + uint64_t(s0) * n;
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane % kPiLanes, lead = lane - sub;
  const uint32_t p = (blockIdx.x * kBlock + threadIdx.x) / kPiLanes;
  uint32_t e0 = 0, e1 = 0, r = 0;
  bool active = false;
  if (p < n) {
    r = jobOfPos[p];
    active = !done[r];
    if (active) {
      e0 = colPtr[p];
      e1 = colPtr[p + 1];
    }
  }
  const double nPrev = active && t >= 2 ? norm[uint64_t(s1) * nR + r] : 1.0;
  const double nPrev2 = active && t >= 3 ? norm[uint64_t(s2) * nR + r] : 1.0;
  const double mu0 = active ? 1.0 / double(lo[r + 1] - lo[r]) : 0.0;
  uint32_t trips = (e1 - e0 + kPiLanes - 1) / kPiLanes;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) trips = max(trips, (uint32_t)__shfl_xor(trips, o, kWave));
  double s = 0.0;
  for (uint32_t q = 0; q < trips; ++q) {
    const uint32_t e = e0 + q * kPiLanes + sub;
    double prod = 0.0;
    if (e < e1) {
      const uint32_t i = rowIdx[e];
      const double m = (t == 1) ? mu0 : yPrev[i] / nPrev;
      prod = m * val[e];
    }
#pragma unroll
    for (int k = 0; k < kPiLanes; ++k) {
      const double pk = __shfl(prod, lead + k, kWave);
      if (e0 + q * kPiLanes + k < e1) s += pk;
    }
  }
  if (active && sub == 0) {
    yCur[p] = s;
    absCol[p] = fabs(s);
    double res = 0.0;
    if (t >= 2) {
      const double a = yPrev[p] / nPrev;
      const double b = (t == 2) ? mu0 : yPrev2[p] / nPrev2;
      res = fabs(a - b);
    }
    resCol[p] = res;
  }
}

// One block per job: norm of y_t and residual of iteration t-1 over the job's columns.
static __global__ __launch_bounds__(kBlock) void batchPiReduceKernel(const double* absCol, const double* resCol,
                                                              const uint32_t* lo, uint32_t nR, double* norm,
                                                              double* resid, int* done, uint32_t* finalIt,
                                                              BatchPiState* st, uint32_t maxIter, double tol) {
  const uint32_t r = blockIdx.x;
  if (st->nDone >= nR) return;
  const uint32_t t = st->tR;
  if (r == 0 && threadIdx.x == 0) st->t = t + 1;
  if (r >= nR || done[r]) return;
  const uint32_t a = lo[r], b = lo[r + 1];
  double ta = 0.0, tr = 0.0;
  for (uint32_t p = a + threadIdx.x; p < b; p += kBlock) {
    ta += absCol[p];
    tr += resCol[p];
  }
  ta = blockReduceSumF64(ta);
  tr = blockReduceSumF64(tr);
  if (threadIdx.x == 0) {
    norm[uint64_t(t % 3) * nR + r] = ta;
    const uint32_t completed = t - 1;
    if (completed >= 1) {
      resid[r] = tr;
      if ((tol >= 0.0 && tr <= tol) || completed >= maxIter) {
        done[r] = 1;
        finalIt[r] = completed;
        atomicAdd(&st->nDone, 1u);
      }
    }
  }
}

static __global__ __launch_bounds__(kBlock) void batchPiFinishKernel(const double* y, uint32_t n, const uint32_t* jobOfPos,
                                                              const uint32_t* lo, uint32_t nR, const double* norm,
                                                              const uint32_t* finalIt, double* mu) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const uint32_t r = jobOfPos[p];
  const uint32_t f = finalIt[r];
  mu[p] = f == 0 ? 1.0 / double(lo[r + 1] - lo[r]) : y[uint64_t(f % 3) * n + p] / norm[uint64_t(f % 3) * nR + r];
}

static __global__ __launch_bounds__(kBlock) void jobOfPosKernel(const uint32_t* lo, uint32_t nR, uint32_t* jobOfPos) {
  const uint32_t r = blockIdx.x;
  if (r >= nR) return;
  for (uint32_t p = lo[r] + threadIdx.x; p < lo[r + 1]; p += kBlock) jobOfPos[p] = r;
}

}  // namespace b12x
