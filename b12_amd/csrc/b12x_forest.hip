// Batched and served forms of the hot path (b12x_forest_* and b12x_serve_run in
// include/b12x_c_api.h). A forest holds the trees of many jobs in one set of leaf arrays; every
// step runs one subdivision step (subdivision.m:18-25) for all of its active jobs in shared
// launches, and jobs that reach their depth retire through one batched transition matrix
// (ImplicitBoxTree.hpp:2932-3159) and one block-diagonal power iteration with per-job stopping.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/b12x_c_api.h"
#include "b12x_forest_kernels.hpp"
#include "b12x_host.hpp"
#include "b12x_launch.hpp"

using namespace b12x;

namespace {

// Device results of one retirement wave; jobs keep offsets into it.
struct ResultArena {
  uint64_t* key = nullptr;
  uint8_t* depth = nullptr;
  uint8_t* flags = nullptr;
  uint32_t* rows = nullptr;
  uint32_t* cols = nullptr;
  double* vals = nullptr;
  double* mu = nullptr;
  uint64_t bytes = 0;
  ~ResultArena() {
    dfree(key); dfree(depth); dfree(flags); dfree(rows); dfree(cols); dfree(vals); dfree(mu);
  }
};

struct Job {
  int state = 0;  // 0 waiting for upload, 1 active, 4 measure iterating, 2 done, 3 released
  uint8_t target = 0, level = 0;
  JobDev dev{};
  ForcingStep* dForcing = nullptr;
  uint32_t nLeaves = 0;
  uint64_t nnz = 0;
  int iters = 0;
  double resid = -1.0;
  std::shared_ptr<ResultArena> arena;
  uint64_t leafBase = 0, nnzBase = 0;
  double tDone = 0.0;
};

// Host mirror of one job in the power-iteration pool.
struct PoolJob {
  uint32_t slot;      // forest job slot
  uint32_t lo, hi;    // pool positions
  uint64_t elo, ehi;  // pool entries
  double* muOut;      // where the measure goes (the job's place in its result arena)
};

// Block-diagonal transition matrices of finished jobs still iterating towards their measure.
struct Pool {
  uint32_t n = 0, capN = 0, capJ = 0, nB = 0, capB = 0;
  uint64_t nnz = 0, capE = 0;
  uint32_t* colPtr = nullptr;    // capN + 1
  uint32_t* jobOfPos = nullptr;  // capN
  double* y = nullptr;           // 3 slots of stride capN
  uint32_t* rowIdx = nullptr;    // capE
  uint8_t* cnt = nullptr;        // capE: hit count of each entry (value = table[count])
  double* table = nullptr;       // kPoolTable values count / k
  uint32_t* jobOfBlock = nullptr;  // capB: pool job of each step-kernel block
  double* partAbs = nullptr;       // capB
  double* partRes = nullptr;
  uint64_t* coopPart = nullptr;    // 6 * capB: partial slots of the persistent form
  int* coopErr = nullptr;
  bool coopOff = false;            // a persistent wait passed its limit: batched form from then on
  bool coopInFlight = false;
  PoolJobDev* dJobs = nullptr;  // capJ
  std::vector<PoolJob> jobs;
  hipStream_t stream = nullptr;  // the pool's own stream: its batches overlap the subdivision steps
  bool inFlight = false;         // a batch has been launched and not yet harvested
  double tLaunch = 0.0;
  // captured batch of iterations for the current pool layout (recaptured after joins and exits)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graphIters = 0;
  void dropGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graphIters = 0;
  }
  void freeAll() {
    streamRetire(stream);
    dropGraph();
    dfree(colPtr); dfree(jobOfPos); dfree(y); dfree(rowIdx); dfree(cnt); dfree(table); dfree(jobOfBlock);
    dfree(partAbs); dfree(partRes); dfree(dJobs); dfree(coopPart); dfree(coopErr);
    cnt = nullptr;
    table = nullptr;
    coopPart = nullptr;
    coopErr = nullptr;
    colPtr = jobOfPos = rowIdx = jobOfBlock = nullptr;
    y = partAbs = partRes = nullptr;
    dJobs = nullptr;
    streamRelease(stream);
    stream = nullptr;
  }
};

double nowSeconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr int kForestScal = 16;

}  // namespace

struct b12x_forest_s {
  int dim = 0, mapKind = 0;
  GridParams grid{};
  double* dTable = nullptr;
  int useAbs = 0, rowMajor = 0, piMaxIter = 200;
  double piTol = -1.0;
  hipStream_t stream = nullptr;
  uint32_t n = 0, cap = 0;
  ForestView arr[2]{};
  int cur = 0;
  uint8_t* hit = nullptr;
  uint32_t* totals = nullptr;
  uint32_t* scal = nullptr;
  uint32_t* hscal = nullptr;
  uint32_t slotCap = 0;
  JobDev* dJobs = nullptr;
  uint32_t* dOff = nullptr;
  uint32_t* dEnd = nullptr;
  uint8_t* dRetiring = nullptr;
  std::vector<Job> jobs;
  std::vector<uint32_t> active;   // slots of active jobs
  std::vector<uint32_t> waiting;  // slots added but not yet uploaded
  double imageMs = 0.0;
  uint64_t steps = 0;
  // telemetry: wall seconds per phase, pool iterations, pool high-water marks
  double tSubdiv = 0.0, tRetire = 0.0, tPool = 0.0;
  uint64_t poolIters = 0, poolMaxJobs = 0, poolMaxN = 0, poolMaxNnz = 0;
  double poolColIters = 0.0, poolNnzIters = 0.0;  // sums of pool columns and entries over iterations
  Pool pool;
  int poolItersPerStep = 64;

  ~b12x_forest_s() {
    streamRetire(stream);
    pool.freeAll();
    for (int b = 0; b < 2; ++b) {
      dfree(arr[b].key); dfree(arr[b].depth); dfree(arr[b].flags); dfree(arr[b].job);
    }
    dfree(hit); dfree(totals); dfree(scal); dfree(dTable); dfree(dJobs); dfree(dOff); dfree(dEnd); dfree(dRetiring);
    for (auto& j : jobs) dfree(j.dForcing);
    if (hscal) (void)hipHostFree(hscal);
    streamRelease(stream);
  }
};

namespace {

void fSync(b12x_forest_s& F, int count) {
  HIPCHK(hipMemcpyAsync(F.hscal, F.scal, count * sizeof(uint32_t), hipMemcpyDeviceToHost, F.stream));
  HIPCHK(hipStreamSynchronize(F.stream));
}

void ensureLeafCap(b12x_forest_s& F, uint64_t need) {
  if (need > 0x7FFFFFFEull) throw HipFail(B12X_E_ARG, "forest exceeds 2^31 leaves");
  if (need <= F.cap) return;
  const uint64_t nc = std::max<uint64_t>({need, uint64_t(F.cap) * 2, 4096});
  ForestView a[2];
  for (int b = 0; b < 2; ++b) {
    a[b].key = dalloc<uint64_t>(nc);
    a[b].depth = dalloc<uint8_t>(nc);
    a[b].flags = dalloc<uint8_t>(nc);
    a[b].job = dalloc<uint32_t>(nc);
  }
  if (F.n) {
    const ForestView& c = F.arr[F.cur];
    HIPCHK(hipMemcpyAsync(a[0].key, c.key, size_t(F.n) * 8, hipMemcpyDeviceToDevice, F.stream));
    HIPCHK(hipMemcpyAsync(a[0].depth, c.depth, F.n, hipMemcpyDeviceToDevice, F.stream));
    HIPCHK(hipMemcpyAsync(a[0].flags, c.flags, F.n, hipMemcpyDeviceToDevice, F.stream));
    HIPCHK(hipMemcpyAsync(a[0].job, c.job, size_t(F.n) * 4, hipMemcpyDeviceToDevice, F.stream));
  }
  uint8_t* h = dalloc<uint8_t>(nc);
  HIPCHK(hipMemsetAsync(h, 0, nc, F.stream));
  uint32_t* tt = dalloc<uint32_t>(nc * 2 / kTile + 16);
  HIPCHK(hipStreamSynchronize(F.stream));
  for (int b = 0; b < 2; ++b) {
    dfree(F.arr[b].key); dfree(F.arr[b].depth); dfree(F.arr[b].flags); dfree(F.arr[b].job);
    F.arr[b] = a[b];
  }
  dfree(F.hit);
  dfree(F.totals);
  F.hit = h;
  F.totals = tt;
  F.cur = 0;
  F.cap = uint32_t(nc);
}

void ensureSlotCap(b12x_forest_s& F, uint32_t need) {
  if (need <= F.slotCap) return;
  const uint32_t nc = std::max<uint32_t>({need, F.slotCap * 2, 64});
  JobDev* j = dalloc<JobDev>(nc);
  uint32_t* o = dalloc<uint32_t>(nc);
  uint32_t* e = dalloc<uint32_t>(nc);
  uint8_t* r = dalloc<uint8_t>(nc);
  HIPCHK(hipMemsetAsync(r, 0, nc, F.stream));
  HIPCHK(hipMemsetAsync(o, 0, size_t(nc) * 4, F.stream));
  HIPCHK(hipMemsetAsync(e, 0, size_t(nc) * 4, F.stream));
  if (F.slotCap) {
    HIPCHK(hipMemcpyAsync(j, F.dJobs, size_t(F.slotCap) * sizeof(JobDev), hipMemcpyDeviceToDevice, F.stream));
    HIPCHK(hipMemcpyAsync(o, F.dOff, size_t(F.slotCap) * 4, hipMemcpyDeviceToDevice, F.stream));
    HIPCHK(hipMemcpyAsync(e, F.dEnd, size_t(F.slotCap) * 4, hipMemcpyDeviceToDevice, F.stream));
  }
  HIPCHK(hipStreamSynchronize(F.stream));
  dfree(F.dJobs); dfree(F.dOff); dfree(F.dEnd); dfree(F.dRetiring);
  F.dJobs = j;
  F.dOff = o;
  F.dEnd = e;
  F.dRetiring = r;
  F.slotCap = nc;
}

template <class Op>
void fScan(b12x_forest_s& F, const Op& op, uint32_t n, int scalSlot, const uint32_t* nDev = nullptr) {
  if (n == 0) return;
  const uint32_t nb = ceilDiv(n, kTile);
  hipLaunchKernelGGL(scanCountKernel<Op>, dim3(nb), dim3(kBlock), 0, F.stream, op, n, nDev, F.totals,
                     (uint32_t*)nullptr);
  hipLaunchKernelGGL(scanTotalsKernel, dim3(1), dim3(kBlock), 0, F.stream, F.totals, nb, nDev, F.scal + scalSlot,
                     (const uint32_t*)nullptr, (uint32_t*)nullptr);
  hipLaunchKernelGGL(scanScatterKernel<Op>, dim3(nb), dim3(kBlock), 0, F.stream, op, n, nDev, F.totals);
  HIPCHK(hipGetLastError());
}

// Recompute every job's leaf range from the job ids (n leaves, or *nDev with n a bound).
void fBoundary(b12x_forest_s& F, uint32_t n, const uint32_t* nDev = nullptr) {
  hipLaunchKernelGGL(forestClearRangesKernel, dim3(ceilDiv(F.slotCap, kBlock)), dim3(kBlock), 0, F.stream, F.dOff,
                     F.dEnd, F.slotCap);
  if (n)
    hipLaunchKernelGGL(forestBoundaryKernel, dim3(ceilDiv(n, kBlock)), dim3(kBlock), 0, F.stream,
                       (const uint32_t*)F.arr[F.cur].job, n, nDev, F.dOff, F.dEnd);
  HIPCHK(hipGetLastError());
}

template <int DIM, int MAP>
void launchForestImage2(const ForestImageArgs& a, uint64_t pointsBound, hipStream_t st) {
  const uint32_t blocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((pointsBound + kBlock - 1) / kBlock,
                                                                              1u << 22)));
  if (a.grid.kind == GRID_INNER)
    hipLaunchKernelGGL((forestImageKernel<DIM, MAP, GRID_INNER>), dim3(blocks), dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((forestImageKernel<DIM, MAP, GRID_UNIFORM>), dim3(blocks), dim3(kBlock), 0, st, a);
}

void launchForestImage(int dim, int mapKind, const ForestImageArgs& a, uint64_t pointsBound, hipStream_t st) {
  switch (mapKind) {
    case MAP_IDENTITY:
      if (dim == 1) launchForestImage2<1, MAP_IDENTITY>(a, pointsBound, st);
      else if (dim == 2) launchForestImage2<2, MAP_IDENTITY>(a, pointsBound, st);
      else if (dim == 3) launchForestImage2<3, MAP_IDENTITY>(a, pointsBound, st);
      else launchForestImage2<4, MAP_IDENTITY>(a, pointsBound, st);
      break;
    case MAP_HENON: launchForestImage2<2, MAP_HENON>(a, pointsBound, st); break;
    case MAP_LORENZ: launchForestImage2<3, MAP_LORENZ>(a, pointsBound, st); break;
    case MAP_HENON_HEILES: launchForestImage2<4, MAP_HENON_HEILES>(a, pointsBound, st); break;
    case MAP_DUFFING: launchForestImage2<2, MAP_DUFFING>(a, pointsBound, st); break;
    default: throw HipFail(B12X_E_UNSUPPORTED, "map kind");
  }
  HIPCHK(hipGetLastError());
}

ForestImageArgs imageArgs(b12x_forest_s& F) {
  ForestImageArgs a{};
  a.jobs = F.dJobs;
  a.off = F.dOff;
  a.end = F.dEnd;
  a.key = F.arr[F.cur].key;
  a.depth = F.arr[F.cur].depth;
  a.job = F.arr[F.cur].job;
  a.grid = F.grid;
  return a;
}

// Upload the root leaves of newly added jobs (one root each, flags SD: setFlags(SD, 0) of a new
// tree, subdivision.m:15) and their parameter tables.
void admitWaiting(b12x_forest_s& F) {
  if (F.waiting.empty()) return;
  const uint32_t m = uint32_t(F.waiting.size());
  ensureLeafCap(F, uint64_t(F.n) + m);
  std::vector<uint64_t> k(m, keyOf(0, 0));
  std::vector<uint8_t> d(m, 0), f(m, uint8_t(F_SD));
  std::vector<uint32_t> j(F.waiting.begin(), F.waiting.end());
  const ForestView& c = F.arr[F.cur];
  HIPCHK(hipMemcpyAsync(c.key + F.n, k.data(), size_t(m) * 8, hipMemcpyHostToDevice, F.stream));
  HIPCHK(hipMemcpyAsync(c.depth + F.n, d.data(), m, hipMemcpyHostToDevice, F.stream));
  HIPCHK(hipMemcpyAsync(c.flags + F.n, f.data(), m, hipMemcpyHostToDevice, F.stream));
  HIPCHK(hipMemcpyAsync(c.job + F.n, j.data(), size_t(m) * 4, hipMemcpyHostToDevice, F.stream));
  for (uint32_t s : F.waiting)
    HIPCHK(hipMemcpyAsync(F.dJobs + s, &F.jobs[s].dev, sizeof(JobDev), hipMemcpyHostToDevice, F.stream));
  HIPCHK(hipStreamSynchronize(F.stream));
  F.n += m;
  for (uint32_t s : F.waiting) {
    F.jobs[s].state = 1;
    F.active.push_back(s);
  }
  F.waiting.clear();
  fBoundary(F, F.n);
}

static __global__ __launch_bounds__(kBlock) void poolFillU64Kernel(uint64_t* p, uint64_t n, uint64_t v) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// Largest pool (in step blocks) the persistent form takes: every block must be resident at once.
uint32_t poolCoopBlocks() {
  static uint32_t cached = 0xFFFFFFFFu;
  if (cached != 0xFFFFFFFFu) return cached;
  int dev = 0, cus = 0, occ = 0;
  cached = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(&poolCoopKernel), kBlock, 0) ==
          hipSuccess &&
      occ > 0)
    cached = uint32_t(std::min<uint64_t>(kPoolCoopMaxBlocks, uint64_t(occ) * uint64_t(cus)));
  return cached;
}

// Poll limit of the persistent form's waits (B12X_POOL_SPIN overrides it; 0 makes every wait
// give up at once, which tests use to run the fallback).
uint32_t poolSpinLimit() {
  const char* e = std::getenv("B12X_POOL_SPIN");
  return e ? uint32_t(std::strtoul(e, nullptr, 10)) : (1u << 21);
}

static __global__ __launch_bounds__(kBlock) void narrowCountKernel(const uint32_t* in, uint64_t n, uint8_t* out) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = uint8_t(in[i]);
}

static __global__ __launch_bounds__(kBlock) void addCopyKernel(const uint32_t* in, uint64_t n, uint32_t add,
                                                               uint32_t* out) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = in[i] + add;
}

template <class T>
void growCopy(T*& p, uint64_t used, uint64_t newCap, hipStream_t st) {
  T* q = dalloc<T>(newCap);
  if (used) HIPCHK(hipMemcpyAsync(q, p, used * sizeof(T), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  dfree(p);
  p = q;
}

// Append a wave of finished jobs: wave positions 0..m-1 (job-major), CSC colPtr[m+1] over them,
// rows as wave positions, values, and each job's wave position and entry range.
void poolJoin(b12x_forest_s& F, uint32_t m, uint64_t nnz, const uint32_t* colPtr, const uint32_t* rows,
              const uint32_t* counts, const std::vector<PoolJob>& joiners) {
  Pool& P = F.pool;
  hipStream_t st = P.stream;  // idle here: the last batch has been harvested
  P.dropGraph();
  const uint32_t n0 = P.n;
  const uint64_t e0 = P.nnz;
  if (uint64_t(n0) + m + 1 > 0xFFFFFFFEull || e0 + nnz > 0xFFFFFFFEull)
    throw HipFail(B12X_E_UNSUPPORTED, "power-iteration pool exceeds 2^32 entries");
  if (n0 + m + 1 > P.capN) {
    const uint32_t c = std::max<uint32_t>({n0 + m + 1, P.capN * 2, 4096});
    growCopy(P.colPtr, n0 + (n0 ? 1 : 0), uint64_t(c) + 1, st);
    growCopy(P.jobOfPos, n0, c, st);
    double* y = dalloc<double>(uint64_t(c) * 3);
    for (int sl = 0; sl < 3 && n0; ++sl)
      HIPCHK(hipMemcpyAsync(y + uint64_t(sl) * c, P.y + uint64_t(sl) * P.capN, uint64_t(n0) * 8,
                            hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(P.y);
    P.y = y;
    P.capN = c;
  }
  if (e0 + nnz > P.capE) {
    const uint64_t c = std::max<uint64_t>({e0 + nnz, P.capE * 2, 65536});
    growCopy(P.rowIdx, e0, c, st);
    growCopy(P.cnt, e0, c, st);
    P.capE = c;
  }
  const uint32_t nJ0 = uint32_t(P.jobs.size());
  const uint32_t nJ = nJ0 + uint32_t(joiners.size());
  uint32_t nB = P.nB;
  for (const PoolJob& pj : joiners) nB += poolBlocksOf(pj.hi - pj.lo);
  if (nB > P.capB) {
    const uint32_t c = std::max<uint32_t>({nB, P.capB * 2, 1024});
    growCopy(P.jobOfBlock, P.nB, c, st);
    dfree(P.partAbs);
    dfree(P.partRes);
    dfree(P.coopPart);
    P.partAbs = dalloc<double>(c);
    P.partRes = dalloc<double>(c);
    P.coopPart = dalloc<uint64_t>(uint64_t(c) * 6);
    P.capB = c;
  }
  if (nJ > P.capJ) {
    const uint32_t c = std::max<uint32_t>({nJ, P.capJ * 2, 64});
    growCopy(P.dJobs, nJ0, c, st);
    P.capJ = c;
  }
  // colPtr (m+1 values, rebased) ; rows rebased to pool positions ; values ; job of position
  hipLaunchKernelGGL(addCopyKernel, dim3(ceilDiv(uint64_t(m) + 1, kBlock)), dim3(kBlock), 0, st, colPtr,
                     uint64_t(m) + 1, uint32_t(e0), P.colPtr + n0);
  if (nnz) {
    hipLaunchKernelGGL(addCopyKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, rows, nnz, n0,
                       P.rowIdx + e0);
    hipLaunchKernelGGL(narrowCountKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, counts, nnz, P.cnt + e0);
  }
  std::vector<uint32_t> jop(m);
  std::vector<PoolJobDev> jd(joiners.size());
  uint32_t b = P.nB;
  for (size_t r = 0; r < joiners.size(); ++r) {
    PoolJob pj = joiners[r];
    for (uint32_t q = pj.lo; q < pj.hi; ++q) jop[q] = nJ0 + uint32_t(r);
    pj.lo += n0;
    pj.hi += n0;
    pj.elo += e0;
    pj.ehi += e0;
    jd[r] = PoolJobDev{pj.lo, pj.hi, 1u, 0, 0u, b, {0.0, 0.0, 0.0}, -1.0, 0.0};
    b += poolBlocksOf(pj.hi - pj.lo);
    P.jobs.push_back(pj);
  }
  if (m) HIPCHK(hipMemcpyAsync(P.jobOfPos + n0, jop.data(), uint64_t(m) * 4, hipMemcpyHostToDevice, st));
  if (!jd.empty()) {
    HIPCHK(hipMemcpyAsync(P.dJobs + nJ0, jd.data(), jd.size() * sizeof(PoolJobDev), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(poolJobOfBlockKernel, dim3(uint32_t(jd.size())), dim3(kBlock), 0, st,
                       (const PoolJobDev*)P.dJobs, nJ0, nJ, P.jobOfBlock);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));
  P.n = n0 + m;
  P.nnz = e0 + nnz;
  P.nB = nB;
}

// Launch a batch of `iters` iterations of every pool job on the pool's stream (asynchronous; the
// forest's subdivision work runs meanwhile on its own stream).
void poolLaunch(b12x_forest_s& F, int iters) {
  Pool& P = F.pool;
  if (P.jobs.empty() || P.inFlight) return;
  P.tLaunch = nowSeconds();
  F.poolIters += uint64_t(iters);
  F.poolMaxJobs = std::max<uint64_t>(F.poolMaxJobs, P.jobs.size());
  F.poolMaxN = std::max<uint64_t>(F.poolMaxN, P.n);
  F.poolMaxNnz = std::max<uint64_t>(F.poolMaxNnz, P.nnz);
  F.poolColIters += double(P.n) * double(iters);
  F.poolNnzIters += double(P.nnz) * double(iters);
  hipStream_t st = P.stream;
  // small pools: the persistent form (every block resident; B12X_POOL_PERSISTENT=0: batches).
  // With no subdivision running, a launch may run a longer budget.
  const char* pe = std::getenv("B12X_POOL_PERSISTENT");
  if (!P.coopOff && !(pe && pe[0] == '0') && P.nB <= poolCoopBlocks()) {
    if (!P.coopErr) P.coopErr = dalloc<int>(1);
    const uint64_t nPart = uint64_t(P.nB) * 6;
    hipLaunchKernelGGL(poolFillU64Kernel, dim3(ceilDiv(nPart, kBlock)), dim3(kBlock), 0, st, P.coopPart, nPart,
                       kCoopEmpty);
    HIPCHK(hipMemsetAsync(P.coopErr, 0, sizeof(int), st));
    const uint32_t budget = uint32_t(iters) * (F.active.empty() ? 16u : 1u);
    PoolCoopArgs ca{P.colPtr, P.rowIdx, P.cnt, P.table, P.jobOfBlock, P.dJobs, P.y, P.capN, P.coopPart, P.nB, P.coopErr,
                    uint32_t(std::max(F.piMaxIter, 0)), F.piTol, budget, poolSpinLimit()};
    hipLaunchKernelGGL(poolCoopKernel, dim3(P.nB), dim3(kBlock), 0, st, ca);
    HIPCHK(hipGetLastError());
    F.poolIters += budget - uint64_t(iters);
    F.poolColIters += double(P.n) * double(budget - uint64_t(iters));
    F.poolNnzIters += double(P.nnz) * double(budget - uint64_t(iters));
    P.coopInFlight = true;
    P.inFlight = true;
    return;
  }
  auto enqueue = [&]() {
    for (int b = 0; b < iters; ++b) {
      hipLaunchKernelGGL(poolStepKernel, dim3(P.nB), dim3(kBlock), 0, st, (const uint32_t*)P.colPtr,
                         (const uint32_t*)P.rowIdx, (const uint8_t*)P.cnt, (const double*)P.table,
                         (const uint32_t*)P.jobOfBlock, P.dJobs, P.y, P.capN, P.partAbs, P.partRes,
                         uint32_t(std::max(F.piMaxIter, 0)), F.piTol);
    }
  };
  // the batch runs from a captured HIP graph (B12X_GRAPHS=0: eager launches)
  const char* ge = std::getenv("B12X_GRAPHS");
  if (ge && ge[0] == '0') {
    enqueue();
  } else {
    if (!P.exec || P.graphIters != iters) {
      P.dropGraph();
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue();
      HIPCHK(hipStreamEndCapture(st, &P.graph));
      HIPCHK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
      P.graphIters = iters;
    }
    HIPCHK(hipGraphLaunch(P.exec, st));
  }
  HIPCHK(hipGetLastError());
  P.inFlight = true;
}

// Wait for the launched batch, then hand finished jobs their measures and drop them from the pool.
void poolHarvest(b12x_forest_s& F) {
  Pool& P = F.pool;
  if (!P.inFlight) return;
  P.inFlight = false;
  hipStream_t st = P.stream;
  const uint32_t nJ = uint32_t(P.jobs.size());
  std::vector<PoolJobDev> hj(nJ);
  HIPCHK(hipMemcpyAsync(hj.data(), P.dJobs, nJ * sizeof(PoolJobDev), hipMemcpyDeviceToHost, st));
  int coopErr = 0;
  if (P.coopInFlight) HIPCHK(hipMemcpyAsync(&coopErr, P.coopErr, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (P.coopInFlight && coopErr) {
    // a persistent wait gave up and every block has left: the jobs still running start again
    // from iteration 1 in the batched form (their iterates do not depend on the form)
    P.coopOff = true;
    for (uint32_t r = 0; r < nJ; ++r)
      if (!hj[r].done) {
        hj[r].t = 1;
        hj[r].norm[0] = hj[r].norm[1] = hj[r].norm[2] = 0.0;
        hj[r].resid = -1.0;
        hj[r].nuLag = 0.0;
        hj[r].arrived = 0;
      }
    HIPCHK(hipMemcpyAsync(P.dJobs, hj.data(), nJ * sizeof(PoolJobDev), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  P.coopInFlight = false;
  bool anyDone = false;
  for (uint32_t r = 0; r < nJ; ++r) anyDone = anyDone || hj[r].done;
  if (!anyDone) {
    F.tPool += nowSeconds() - P.tLaunch;
    return;
  }
  P.dropGraph();
  const double tNow = nowSeconds();
  std::vector<int> keep(nJ, 1);
  std::vector<uint32_t> posShift(nJ, 0), nnzShift(nJ, 0), newIndex(nJ, 0);
  std::vector<PoolJob> kept;
  std::vector<PoolJobDev> keptDev;
  uint32_t nNew = 0, bNew = 0;
  uint64_t eNew = 0;
  for (uint32_t r = 0; r < nJ; ++r) {
    PoolJob& pj = P.jobs[r];
    if (hj[r].done) {
      keep[r] = 0;
      hipLaunchKernelGGL(poolFinishKernel, dim3(std::max<uint32_t>(1, ceilDiv(pj.hi - pj.lo, kBlock))), dim3(kBlock), 0, st,
                         (const double*)P.y, P.capN, (const PoolJobDev*)P.dJobs, r, pj.muOut);
      Job& J = F.jobs[pj.slot];
      J.iters = int(hj[r].final);
      J.resid = hj[r].resid;
      J.state = 2;
      J.tDone = tNow;
      continue;
    }
    posShift[r] = pj.lo - nNew;
    nnzShift[r] = uint32_t(pj.elo - eNew);
    newIndex[r] = uint32_t(kept.size());
    PoolJobDev d = hj[r];
    const uint32_t len = pj.hi - pj.lo;
    const uint64_t elen = pj.ehi - pj.elo;
    d.lo = nNew;
    d.hi = nNew + len;
    d.b0 = bNew;
    bNew += poolBlocksOf(len);
    PoolJob q = pj;
    q.lo = nNew;
    q.hi = nNew + len;
    q.elo = eNew;
    q.ehi = eNew + elen;
    kept.push_back(q);
    keptDev.push_back(d);
    nNew += len;
    eNew += elen;
  }
  HIPCHK(hipGetLastError());
  // compact into fresh buffers
  uint32_t* dShiftP = dalloc<uint32_t>(nJ);
  uint32_t* dShiftE = dalloc<uint32_t>(nJ);
  uint32_t* dNewIdx = dalloc<uint32_t>(nJ);
  int* dKeep = dalloc<int>(nJ);
  HIPCHK(hipMemcpyAsync(dShiftP, posShift.data(), nJ * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dShiftE, nnzShift.data(), nJ * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dNewIdx, newIndex.data(), nJ * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dKeep, keep.data(), nJ * sizeof(int), hipMemcpyHostToDevice, st));
  uint32_t* colPtr = dalloc<uint32_t>(uint64_t(P.capN) + 1);
  uint32_t* jobOfPos = dalloc<uint32_t>(P.capN);
  double* y = dalloc<double>(uint64_t(P.capN) * 3);
  uint32_t* rowIdx = dalloc<uint32_t>(P.capE);
  uint8_t* cnts = dalloc<uint8_t>(P.capE);
  if (P.n)
    hipLaunchKernelGGL(poolMoveColumnsKernel, dim3(ceilDiv(P.n, kBlock)), dim3(kBlock), 0, st,
                       (const uint32_t*)P.colPtr, (const uint32_t*)P.jobOfPos, (const double*)P.y, P.n, P.capN,
                       (const uint32_t*)dShiftP, (const uint32_t*)dShiftE, (const int*)dKeep,
                       (const uint32_t*)dNewIdx, colPtr, jobOfPos, y, P.capN);
  if (P.nnz)
    hipLaunchKernelGGL(poolMoveEntriesKernel, dim3(ceilDiv(P.nnz, kBlock)), dim3(kBlock), 0, st,
                       (const uint32_t*)P.rowIdx, (const uint8_t*)P.cnt, P.nnz, (const uint32_t*)P.jobOfPos,
                       (const uint32_t*)dShiftP, (const uint32_t*)dShiftE, (const int*)dKeep, rowIdx, cnts);
  HIPCHK(hipGetLastError());
  const uint32_t eEnd = uint32_t(eNew);
  HIPCHK(hipMemcpyAsync(colPtr + nNew, &eEnd, 4, hipMemcpyHostToDevice, st));
  if (!keptDev.empty()) {
    HIPCHK(hipMemcpyAsync(P.dJobs, keptDev.data(), keptDev.size() * sizeof(PoolJobDev), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(poolJobOfBlockKernel, dim3(uint32_t(keptDev.size())), dim3(kBlock), 0, st,
                       (const PoolJobDev*)P.dJobs, 0u, uint32_t(keptDev.size()), P.jobOfBlock);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));
  dfree(P.colPtr); dfree(P.jobOfPos); dfree(P.y); dfree(P.rowIdx); dfree(P.cnt);
  P.colPtr = colPtr;
  P.jobOfPos = jobOfPos;
  P.y = y;
  P.rowIdx = rowIdx;
  P.cnt = cnts;
  P.jobs.swap(kept);
  P.n = nNew;
  P.nnz = eNew;
  P.nB = bNew;
  dfree(dShiftP); dfree(dShiftE); dfree(dNewIdx); dfree(dKeep);
  F.tPool += nowSeconds() - P.tLaunch;
}

// Retire the jobs in R (slots): batched transition matrix, copy of the covers, entry into the
// power-iteration pool, then removal of their leaves from the forest.
void retire(b12x_forest_s& F, std::vector<uint32_t> R, const std::vector<uint32_t>& off,
            const std::vector<uint32_t>& end) {
  std::sort(R.begin(), R.end(), [&](uint32_t a, uint32_t b) { return off[a] < off[b]; });
  const uint32_t nR = uint32_t(R.size());
  std::vector<uint32_t> lo(nR + 1, 0), offR(nR);
  for (uint32_t r = 0; r < nR; ++r) {
    lo[r + 1] = lo[r] + (end[R[r]] - off[R[r]]);
    offR[r] = off[R[r]];
  }
  const uint32_t nList = lo[nR];
  const uint32_t k = F.grid.k;
  hipStream_t st = F.stream;
  // retiring mask
  std::vector<uint8_t> mask(F.slotCap, 0);
  for (uint32_t s : R) mask[s] = 1;
  HIPCHK(hipMemcpyAsync(F.dRetiring, mask.data(), F.slotCap, hipMemcpyHostToDevice, st));
  // source list (array order = job-major)
  uint32_t* list = dalloc<uint32_t>(std::max<uint32_t>(F.n, nList));  // the scan writes at most F.n entries
  ForestListOp lop{};
  lop.job = F.arr[F.cur].job;
  lop.retiring = F.dRetiring;
  lop.out = list;
  fScan(F, lop, F.n, 0);
  fSync(F, 1);
  if (F.hscal[0] != nList) throw HipFail(B12X_E_HIP, "forest leaf ranges disagree with the job ids");
  // images of every test point of the retiring jobs
  const uint64_t P = uint64_t(nList) * k;
  uint32_t* col = dalloc<uint32_t>(P);
  ForestImageArgs a = imageArgs(F);
  a.srcList = list;
  a.nSrc = nList;
  a.mode = 1;
  a.out = col;
  launchForestImage(F.dim, F.mapKind, a, P, st);
  // rows: per-box sort + run-length encoding (invalid targets dropped)
  uint32_t* rowPtr = dalloc<uint32_t>(uint64_t(nList) + 1);
  TmRowArgs ta{};
  ta.col = col;
  ta.nRows = nList;
  ta.k = k;
  ta.keepInvalid = 0;
  ta.rowCount = rowPtr;
  ta.anyInvalid = F.scal + 4;
  ta.pass = 0;
  tmRows(ta, st);
  HIPCHK(hipMemsetAsync(rowPtr + nList, 0, 4, st));
  const uint32_t nnz = scanArray(st, rowPtr, uint64_t(nList) + 1, F.totals, F.scal + 5, F.hscal + 5);
  uint32_t* ecol = dalloc<uint32_t>(nnz);
  uint32_t* ecnt = dalloc<uint32_t>(nnz);
  uint32_t* erow = dalloc<uint32_t>(nnz);
  ta.pass = 1;
  ta.rowPtr = rowPtr;
  ta.outCol = ecol;
  ta.outCnt = ecnt;
  tmRows(ta, st);
  if (nList)
    hipLaunchKernelGGL(expandRowsKernel, dim3(ceilDiv(nList, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)rowPtr,
                       uint64_t(nList), 0u, erow);
  HIPCHK(hipGetLastError());
  std::vector<uint32_t> hRowPtr(uint64_t(nList) + 1);
  HIPCHK(hipMemcpyAsync(hRowPtr.data(), rowPtr, (uint64_t(nList) + 1) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  dfree(col);
  // column-major copy (stable radix sort by global column leaves rows ascending per column)
  uint32_t* ccol = dalloc<uint32_t>(nnz);
  uint32_t* crow = dalloc<uint32_t>(nnz);
  uint32_t* ccnt = dalloc<uint32_t>(nnz);
  if (nnz) {
    HIPCHK(hipMemcpyAsync(ccol, ecol, uint64_t(nnz) * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(crow, erow, uint64_t(nnz) * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ccnt, ecnt, uint64_t(nnz) * 4, hipMemcpyDeviceToDevice, st));
  }
  radixSort3(st, ccol, crow, ccnt, nnz, bitsFor(uint64_t(F.n) + 1), F.hscal + 6);
  // local indices and columns as list positions
  uint32_t* dLo = dalloc<uint32_t>(nR + 1);
  uint32_t* dOffR = dalloc<uint32_t>(nR);
  uint32_t* jobOfPos = dalloc<uint32_t>(nList);
  HIPCHK(hipMemcpyAsync(dLo, lo.data(), (nR + 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dOffR, offR.data(), nR * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(jobOfPosKernel, dim3(nR), dim3(kBlock), 0, st, (const uint32_t*)dLo, nR, jobOfPos);
  uint32_t* cLocRow = dalloc<uint32_t>(nnz);
  uint32_t* cLocCol = dalloc<uint32_t>(nnz);
  uint32_t* cColPos = dalloc<uint32_t>(nnz);
  uint32_t* rLocRow = dalloc<uint32_t>(nnz);
  uint32_t* rLocCol = dalloc<uint32_t>(nnz);
  uint32_t* rColPos = dalloc<uint32_t>(nnz);
  if (nnz) {
    hipLaunchKernelGGL(forestLocalizeKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)crow,
                       (const uint32_t*)ccol, uint64_t(nnz), (const uint32_t*)jobOfPos, (const uint32_t*)dLo,
                       (const uint32_t*)dOffR, cLocRow, cLocCol, cColPos);
    hipLaunchKernelGGL(forestLocalizeKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)erow,
                       (const uint32_t*)ecol, uint64_t(nnz), (const uint32_t*)jobOfPos, (const uint32_t*)dLo,
                       (const uint32_t*)dOffR, rLocRow, rLocCol, rColPos);
  }
  double* cval = dalloc<double>(nnz);
  double* rval = dalloc<double>(nnz);
  if (nnz) {
    hipLaunchKernelGGL(valuesKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)ccnt,
                       uint64_t(nnz), double(k), F.useAbs, cval);
    hipLaunchKernelGGL(valuesKernel, dim3(ceilDiv(nnz, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)ecnt,
                       uint64_t(nnz), double(k), F.useAbs, rval);
  }
  HIPCHK(hipGetLastError());
  // CSC over list positions and the per-job entry ranges of the column-major order
  uint32_t* colPtr = dalloc<uint32_t>(uint64_t(nList) + 1);
  hipLaunchKernelGGL(compressKernel, dim3(ceilDiv(uint64_t(nList) + 1, kBlock)), dim3(kBlock), 0, st,
                     (const uint32_t*)cColPos, uint64_t(nnz), nList, colPtr);
  std::vector<uint32_t> hColPtr(uint64_t(nList) + 1);
  HIPCHK(hipMemcpyAsync(hColPtr.data(), colPtr, (uint64_t(nList) + 1) * 4, hipMemcpyDeviceToHost, st));
  // results arena: covers, matrices (requested order), measures
  auto A = std::make_shared<ResultArena>();
  A->key = dalloc<uint64_t>(nList);
  A->depth = dalloc<uint8_t>(nList);
  A->flags = dalloc<uint8_t>(nList);
  A->rows = dalloc<uint32_t>(nnz);
  A->cols = dalloc<uint32_t>(nnz);
  A->vals = dalloc<double>(nnz);
  A->mu = dalloc<double>(nList);
  A->bytes = uint64_t(nList) * (8 + 1 + 1 + 8) + uint64_t(nnz) * 16;
  if (nnz) {
    HIPCHK(hipMemcpyAsync(A->rows, F.rowMajor ? rLocRow : cLocRow, uint64_t(nnz) * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(A->cols, F.rowMajor ? rLocCol : cLocCol, uint64_t(nnz) * 4, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(A->vals, F.rowMajor ? rval : cval, uint64_t(nnz) * 8, hipMemcpyDeviceToDevice, st));
  }
  const ForestView& c = F.arr[F.cur];
  for (uint32_t r = 0; r < nR; ++r) {
    const uint32_t cnt = lo[r + 1] - lo[r];
    HIPCHK(hipMemcpyAsync(A->key + lo[r], c.key + offR[r], uint64_t(cnt) * 8, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(A->depth + lo[r], c.depth + offR[r], cnt, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(A->flags + lo[r], c.flags + offR[r], cnt, hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  std::vector<PoolJob> joiners(nR);
  for (uint32_t r = 0; r < nR; ++r) {
    Job& J = F.jobs[R[r]];
    J.arena = A;
    J.leafBase = lo[r];
    J.nLeaves = lo[r + 1] - lo[r];
    const uint64_t e0 = F.rowMajor ? hRowPtr[lo[r]] : hColPtr[lo[r]];
    const uint64_t e1 = F.rowMajor ? hRowPtr[lo[r + 1]] : hColPtr[lo[r + 1]];
    J.nnzBase = e0;
    J.nnz = e1 - e0;
    J.state = 4;  // measuring
    joiners[r] = PoolJob{R[r], lo[r], lo[r + 1], hColPtr[lo[r]], hColPtr[lo[r + 1]], A->mu + lo[r]};
  }
  // the measure is iterated in the pool, on probabilities count / k even when useAbs is set
  poolJoin(F, nList, nnz, colPtr, crow, ccnt, joiners);
  dfree(list); dfree(rowPtr); dfree(ecol); dfree(ecnt); dfree(erow); dfree(ccol); dfree(crow); dfree(ccnt);
  dfree(dLo); dfree(dOffR); dfree(jobOfPos); dfree(cLocRow); dfree(cLocCol); dfree(cColPos); dfree(rLocRow);
  dfree(rLocCol); dfree(rColPos); dfree(cval); dfree(rval); dfree(colPtr);
  // drop the retired leaves
  ForestKeepOp kop{};
  kop.in = F.arr[F.cur];
  kop.out = F.arr[1 - F.cur];
  kop.hit = F.hit;
  kop.retiring = F.dRetiring;
  kop.mode = 1;
  fScan(F, kop, F.n, 0);
  F.cur = 1 - F.cur;
  fSync(F, 1);
  F.n = F.hscal[0];
  std::vector<uint8_t> zero(F.slotCap, 0);
  HIPCHK(hipMemcpyAsync(F.dRetiring, zero.data(), F.slotCap, hipMemcpyHostToDevice, st));
  fBoundary(F, F.n);
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint32_t> still;
  for (uint32_t s : F.active)
    if (F.jobs[s].state == 1) still.push_back(s);
  F.active.swap(still);
}

// One subdivision step for every active job, then retirement of the jobs that reached their depth.
void forestStep(b12x_forest_s& F, uint32_t* nRetired) {
  admitWaiting(F);
  if (nRetired) *nRetired = 0;
  poolLaunch(F, F.poolItersPerStep);
  if (F.active.empty()) {
    poolHarvest(F);
    return;
  }
  const double t0 = nowSeconds();
  const uint32_t n = F.n;
  ensureLeafCap(F, uint64_t(n) * 2);
  hipStream_t st = F.stream;
  // subdivide(SD)
  ForestSubdivideOp sop{};
  sop.in = F.arr[F.cur];
  sop.out = F.arr[1 - F.cur];
  sop.mask = F_SD;
  fScan(F, sop, n, 0);
  F.cur = 1 - F.cur;
  fBoundary(F, 2 * n, F.scal + 0);
  // setFlagsByIterators(OverLeaves, HIT): images of every test point of every job
  ForestImageArgs a = imageArgs(F);
  a.nSrc = 2 * n;
  a.nSrcDev = F.scal + 0;
  a.mode = 0;
  a.hit = F.hit;
  launchForestImage(F.dim, F.mapKind, a, uint64_t(2 * n) * F.grid.k, st);
  // remove(HIT) + unsetFlags(HIT)
  ForestKeepOp kop{};
  kop.in = F.arr[F.cur];
  kop.out = F.arr[1 - F.cur];
  kop.hit = F.hit;
  kop.retiring = F.dRetiring;
  kop.mode = 0;
  fScan(F, kop, 2 * n, 2, F.scal + 0);
  F.cur = 1 - F.cur;
  fBoundary(F, 2 * n, F.scal + 2);
  fSync(F, 3);
  F.n = F.hscal[2];
  ++F.steps;
  std::vector<uint32_t> off(F.slotCap), end(F.slotCap);
  HIPCHK(hipMemcpyAsync(off.data(), F.dOff, size_t(F.slotCap) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(end.data(), F.dEnd, size_t(F.slotCap) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // a job whose cover emptied gets its root back with no flags (ImplicitBoxTree.hpp:2363)
  std::vector<uint32_t> empty;
  for (uint32_t s : F.active)
    if (end[s] <= off[s]) empty.push_back(s);
  if (!empty.empty()) {
    ensureLeafCap(F, uint64_t(F.n) + empty.size());
    const ForestView& c = F.arr[F.cur];
    for (uint32_t s : empty) {
      const uint64_t rk = keyOf(0, 0);
      const uint8_t z = 0;
      HIPCHK(hipMemcpyAsync(c.key + F.n, &rk, 8, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(c.depth + F.n, &z, 1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(c.flags + F.n, &z, 1, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(c.job + F.n, &s, 4, hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
      off[s] = F.n;
      end[s] = F.n + 1;
      ++F.n;
    }
    fBoundary(F, F.n);
  }
  std::vector<uint32_t> R;
  for (uint32_t s : F.active) {
    Job& J = F.jobs[s];
    J.level += 1;
    if (J.level >= J.target) R.push_back(s);
  }
  const double t1 = nowSeconds();
  F.tSubdiv += t1 - t0;
  poolHarvest(F);  // the pool is idle before retiring jobs join it
  const double t2 = nowSeconds();
  if (!R.empty()) retire(F, R, off, end);
  F.tRetire += nowSeconds() - t2;
  if (nRetired) *nRetired = uint32_t(R.size());
}

bool forestBusy(const b12x_forest_s& F) { return !F.active.empty() || !F.waiting.empty() || !F.pool.jobs.empty(); }

Job& jobAt(b12x_forest_s& F, uint32_t id) {
  if (id >= F.jobs.size()) throw HipFail(B12X_E_ARG, "no such job");
  return F.jobs[id];
}

}  // namespace

extern "C" {

int b12x_forest_create(int dim, int gk, const uint64_t* gp, int ngp, int mk, int useAbs, int rowMajor, int piMaxIter,
                       double piTol, b12x_forest* out) {
  return guarded([&]() -> int {
    if (!out || dim < 1 || dim > kMaxDim) return setErr(B12X_E_ARG, "bad argument");
    if (int e = ensureDevice()) return e;
    auto* F = new b12x_forest_s();
    try {
      F->dim = dim;
      F->mapKind = mk;
      std::vector<double> tab;
      buildGrid(dim, gk, gp, ngp, F->grid, tab);
      if (F->grid.k > 64) throw HipFail(B12X_E_UNSUPPORTED, "forests support up to 64 points per box");
      if (!tab.empty()) {
        F->dTable = dalloc<double>(tab.size());
        HIPCHK(hipMemcpy(F->dTable, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
        F->grid.table = F->dTable;
      }
      // validate the map kind against the dimension
      MapParams probe;
      std::vector<ForcingStep> ft;
      buildMapParams(dim, mk, nullptr, 0, probe, ft);
      F->useAbs = useAbs;
      F->rowMajor = rowMajor;
      F->piMaxIter = piMaxIter;
      F->piTol = piTol;
      streamCreate(&F->stream);
      streamCreate(&F->pool.stream);
      std::vector<double> vt(kPoolTable);
      for (uint32_t c = 0; c < kPoolTable; ++c) vt[c] = double(c) / double(F->grid.k);
      F->pool.table = dalloc<double>(kPoolTable);
      HIPCHK(hipMemcpy(F->pool.table, vt.data(), kPoolTable * sizeof(double), hipMemcpyHostToDevice));
      F->scal = dalloc<uint32_t>(kForestScal);
      HIPCHK(hipHostMalloc((void**)&F->hscal, kForestScal * sizeof(uint32_t), hipHostMallocDefault));
      ensureSlotCap(*F, 64);
      ensureLeafCap(*F, 4096);
    } catch (...) {
      delete F;
      throw;
    }
    *out = F;
    return B12X_OK;
  });
}

int b12x_forest_destroy(b12x_forest f) {
  delete f;
  return B12X_OK;
}

int b12x_forest_add(b12x_forest f, const double* center, const double* radius, const double* mp, int nmp,
                    uint8_t target, uint32_t* job) {
  return guarded([&]() -> int {
    if (!f || !center || !radius || !job) return setErr(B12X_E_ARG, "null argument");
    if (target < 1 || target > kMaxDepth) return setErr(B12X_E_ARG, "target depth must be 1..64");
    b12x_forest_s& F = *f;
    Job J;
    J.target = target;
    if (!buildDomain(F.dim, center, radius, nullptr, 0, J.dev.dom)) return setErr(B12X_E_ARG, "domain");
    std::vector<ForcingStep> tab;
    buildMapParams(F.dim, F.mapKind, mp, nmp, J.dev.map, tab);
    if (!tab.empty()) {
      J.dForcing = dalloc<ForcingStep>(tab.size());
      HIPCHK(hipMemcpy(J.dForcing, tab.data(), tab.size() * sizeof(ForcingStep), hipMemcpyHostToDevice));
      J.dev.map.forcing = J.dForcing;
    }
    const uint32_t slot = uint32_t(F.jobs.size());
    ensureSlotCap(F, slot + 1);
    F.jobs.push_back(std::move(J));
    F.waiting.push_back(slot);
    *job = slot;
    return B12X_OK;
  });
}

int b12x_forest_step(b12x_forest f, uint32_t* nActive, uint32_t* nRetired) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    forestStep(*f, nRetired);
    if (nActive) *nActive = uint32_t(f->active.size() + f->waiting.size() + f->pool.jobs.size());
    return B12X_OK;
  });
}

int b12x_forest_run(b12x_forest f) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    while (forestBusy(*f)) forestStep(*f, nullptr);
    return B12X_OK;
  });
}

int b12x_forest_job_info(b12x_forest f, uint32_t id, int* state, uint32_t* nLeaves, uint64_t* nnz, int* iters,
                         double* resid) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    Job& J = jobAt(*f, id);
    if (state) *state = J.state;
    if (nLeaves) *nLeaves = J.nLeaves;
    if (nnz) *nnz = J.nnz;
    if (iters) *iters = J.iters;
    if (resid) *resid = J.resid;
    return B12X_OK;
  });
}

int b12x_forest_job_cover(b12x_forest f, uint32_t id, uint64_t* bits, uint8_t* depth, uint8_t* flags) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    Job& J = jobAt(*f, id);
    if (J.state != 2) return setErr(B12X_E_ARG, "job not finished");
    std::vector<uint64_t> k(J.nLeaves);
    std::vector<uint8_t> d(J.nLeaves);
    HIPCHK(hipMemcpy(k.data(), J.arena->key + J.leafBase, uint64_t(J.nLeaves) * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(d.data(), J.arena->depth + J.leafBase, J.nLeaves, hipMemcpyDeviceToHost));
    if (flags) HIPCHK(hipMemcpy(flags, J.arena->flags + J.leafBase, J.nLeaves, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < J.nLeaves; ++i) {
      if (bits) bits[i] = bitsOf(k[i], d[i]);
      if (depth) depth[i] = d[i];
    }
    return B12X_OK;
  });
}

int b12x_forest_job_matrix(b12x_forest f, uint32_t id, uint32_t* rows, uint32_t* cols, double* vals) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    Job& J = jobAt(*f, id);
    if (J.state != 2) return setErr(B12X_E_ARG, "job not finished");
    if (J.nnz) {
      if (rows) HIPCHK(hipMemcpy(rows, J.arena->rows + J.nnzBase, J.nnz * 4, hipMemcpyDeviceToHost));
      if (cols) HIPCHK(hipMemcpy(cols, J.arena->cols + J.nnzBase, J.nnz * 4, hipMemcpyDeviceToHost));
      if (vals) HIPCHK(hipMemcpy(vals, J.arena->vals + J.nnzBase, J.nnz * 8, hipMemcpyDeviceToHost));
    }
    return B12X_OK;
  });
}

int b12x_forest_job_measure(b12x_forest f, uint32_t id, double* mu) {
  return guarded([&]() -> int {
    if (!f || !mu) return setErr(B12X_E_ARG, "null argument");
    Job& J = jobAt(*f, id);
    if (J.state != 2) return setErr(B12X_E_ARG, "job not finished");
    HIPCHK(hipMemcpy(mu, J.arena->mu + J.leafBase, uint64_t(J.nLeaves) * 8, hipMemcpyDeviceToHost));
    return B12X_OK;
  });
}

int b12x_forest_job_release(b12x_forest f, uint32_t id) {
  return guarded([&]() -> int {
    if (!f) return setErr(B12X_E_ARG, "null forest");
    Job& J = jobAt(*f, id);
    if (J.state == 2) {
      J.arena.reset();
      J.state = 3;
    }
    return B12X_OK;
  });
}

int b12x_forest_footprint(b12x_forest f, uint64_t* bytes) {
  return guarded([&]() -> int {
    if (!f || !bytes) return setErr(B12X_E_ARG, "null argument");
    const b12x_forest_s& F = *f;
    uint64_t b = uint64_t(F.cap) * 2 * (8 + 1 + 1 + 4) + F.cap + uint64_t(F.slotCap) * (sizeof(JobDev) + 9);
    b += (uint64_t(F.pool.capN) + 1) * 4 + uint64_t(F.pool.capN) * (4 + 3 * 8) + F.pool.capE * 5 +
         uint64_t(F.pool.capB) * (4 + 2 * 8 + 6 * 8) + uint64_t(F.pool.capJ) * sizeof(PoolJobDev);
    std::vector<const ResultArena*> seen;
    for (const Job& J : F.jobs)
      if (J.arena && std::find(seen.begin(), seen.end(), J.arena.get()) == seen.end()) {
        seen.push_back(J.arena.get());
        b += J.arena->bytes;
      }
    *bytes = b;
    return B12X_OK;
  });
}

int b12x_forest_stats(b12x_forest f, double* out, int n) {
  return guarded([&]() -> int {
    if (!f || !out) return setErr(B12X_E_ARG, "null argument");
    const b12x_forest_s& F = *f;
    const double v[10] = {double(F.steps),     F.tSubdiv,           F.tRetire,          F.tPool,
                          double(F.poolIters), double(F.poolMaxJobs), double(F.poolMaxN), double(F.poolMaxNnz),
                          F.poolColIters,      F.poolNnzIters};
    for (int i = 0; i < n && i < 10; ++i) out[i] = v[i];
    return B12X_OK;
  });
}

int b12x_forest_set_pool_iterations(b12x_forest f, int iters) {
  return guarded([&]() -> int {
    if (!f || iters < 1) return setErr(B12X_E_ARG, "bad argument");
    f->poolItersPerStep = iters;
    return B12X_OK;
  });
}

// Continuous-batching request loop. Requests (sorted by arrival time, seconds from the start)
// go to forest forestOf[i]; a request is admitted once it has arrived and the forests'
// footprint plus reservePerJob stays within budgetBytes. Every pass runs one merged step of
// each forest with work; completion time = end of the step in which the job's measure converged.
// Admission's cost model of one job (bytes): per leaf the two leaf buffers of the forest with their
// doubling slack (2 x 4 x 14 B), the hit byte, the result arena's cover and measure (18 B), the
// pool's column with its slack (64 B) and the transition matrix's transients (4 B per test point
// and 12 B per leaf); per matrix entry the arena (16 B), the pool with slack (10 B) and the
// transients of the batched matrix (48 B).
static double jobCostBytes(uint64_t leaves, uint64_t nnz, uint64_t k) {
  return double(leaves) * double(208 + 4 * k) + double(nnz) * 74.0;
}

int b12x_serve_run(b12x_forest* forests, int nForests, uint32_t nReq, const double* arrival, const int* forestOf,
                   const double* centers, const double* radii, int stride, const double* mapParams, int nmp,
                   const uint8_t* depths, double budgetBytes, double reservePerJob, double* latency,
                   double* completion, uint32_t* jobIds, double* admittedAt, double* stats) {
  return guarded([&]() -> int {
    if (!forests || nForests <= 0 ||
        (nReq && (!arrival || !forestOf || !centers || !radii || !depths || !latency || !completion)))
      return setErr(B12X_E_ARG, "null argument");
    for (uint32_t i = 0; i < nReq; ++i) {
      completion[i] = latency[i] = -1.0;
      if (admittedAt) admittedAt[i] = -1.0;
    }
    for (uint32_t i = 1; i < nReq; ++i)
      if (arrival[i] < arrival[i - 1]) return setErr(B12X_E_ARG, "requests must be sorted by arrival");
    for (uint32_t i = 0; i < nReq; ++i)
      if (forestOf[i] < 0 || forestOf[i] >= nForests) return setErr(B12X_E_ARG, "bad forest index");
    std::vector<std::vector<int>> reqOf(nForests);
    // learned peak cost per (forest, target depth) of finished jobs
    std::vector<std::vector<double>> learned(nForests, std::vector<double>(256, -1.0));
    auto project = [&](int q, int d) {
      if (learned[q][d] >= 0.0) return learned[q][d];
      for (int e = d - 1; e >= 0; --e)  // the deepest smaller depth, doubling per level
        if (learned[q][e] >= 0.0) return learned[q][e] * std::ldexp(1.0, d - e);
      return reservePerJob;
    };
    std::vector<double> projOf(nReq, 0.0);
    double projInFlight = 0.0, peakUsed = 0.0;
    uint32_t next = 0, completed = 0, inFlight = 0, maxInFlight = 0, deferred = 0;
    auto footprint = [&]() {
      uint64_t used = 0;
      for (int q = 0; q < nForests; ++q) {
        uint64_t b = 0;
        b12x_forest_footprint(forests[q], &b);
        used += b;
      }
      return double(used);
    };
    const double t0 = nowSeconds();
    while (completed < nReq) {
      double now = nowSeconds() - t0;
      // admission by arrival and projected HBM: the forests' bytes now (unreleased results
      // included) plus the projected peak of every job in flight and of the new one
      bool blocked = false;
      while (next < nReq && arrival[next] <= now) {
        const int q = forestOf[next];
        const double pj = project(q, depths[next]);
        const double used = footprint();
        peakUsed = std::max(peakUsed, used);
        if (inFlight > 0 && used + projInFlight + pj > budgetBytes) {
          if (!blocked) ++deferred;
          blocked = true;
          break;
        }
        uint32_t id = 0;
        int rc = b12x_forest_add(forests[q], centers + size_t(next) * stride, radii + size_t(next) * stride,
                                 mapParams ? mapParams + size_t(next) * nmp : nullptr, mapParams ? nmp : 0,
                                 depths[next], &id);
        if (rc) return rc;
        if (reqOf[q].size() <= id) reqOf[q].resize(id + 1, -1);
        reqOf[q][id] = int(next);
        if (jobIds) jobIds[next] = id;
        if (admittedAt) admittedAt[next] = now;
        projOf[next] = pj;
        projInFlight += pj;
        ++next;
        ++inFlight;
        maxInFlight = std::max(maxInFlight, inFlight);
      }
      bool work = false;
      for (int q = 0; q < nForests; ++q) {
        b12x_forest_s& F = *forests[q];
        if (!forestBusy(F)) continue;
        work = true;
        forestStep(F, nullptr);
      }
      peakUsed = std::max(peakUsed, footprint());
      const double tEnd = nowSeconds() - t0;
      for (int q = 0; q < nForests; ++q) {
        b12x_forest_s& F = *forests[q];
        for (uint32_t id = 0; id < reqOf[q].size(); ++id) {
          const int r = reqOf[q][id];
          if (r < 0 || completion[r] >= 0.0 || (F.jobs[id].state != 2 && F.jobs[id].state != 3)) continue;
          completion[r] = tEnd;
          latency[r] = tEnd - arrival[r];
          const Job& J = F.jobs[id];
          const double cost = jobCostBytes(J.nLeaves, J.nnz, F.grid.k);
          learned[q][depths[r]] = std::max(learned[q][depths[r]], cost);
          projInFlight -= projOf[r];
          ++completed;
          --inFlight;
        }
      }
      if (!work && next < nReq) {
        const double wait = arrival[next] - (nowSeconds() - t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
      }
    }
    if (stats) {
      stats[0] = peakUsed;
      stats[1] = double(maxInFlight);
      stats[2] = double(deferred);
    }
    return B12X_OK;
  });
}

}  // extern "C"
