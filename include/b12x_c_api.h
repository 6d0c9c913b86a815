/*
 * b12x C ABI — MI355X (gfx950) implementation of B12's set-oriented hot path.
 *
 * Every entry point takes plain pointers and sizes and returns an int status: 0 on success, a
 * negative B12X_E_* code on failure (b12x_last_error() then holds a message). Reference
 * citations are relative to B12c/src/include/ of SON-GPU-Enabler/B12.
 *
 * Object model (ImplicitBoxTree.h:32-640, CooMatrix.h:56-121):
 *   b12x_tree    an implicit box tree: sorted leaves (key, depth, flags) resident in HBM
 *   b12x_scheme  a (Grid, Map) pair — the functors a point-iterator pair is built from
 *                (ImplicitBoxTree.h:266-430, MEX scheme{1,2}: ...mex.hpp.in:1501-1513,1619-1631)
 *   b12x_coo     a transition matrix in COO form resident in HBM (CooMatrix.h:56-121)
 * The point iterator of a call is (scheme, depth_it, flags_to_check): depth_it = 255 for
 * makePointIteratorPairOverLeaves / ByFlags(-1, flags) and 0 for makePointIteratorPairOverRoot.
 * Depth arguments follow B12: 255 (= -1) is the leaf depth, 0 the root, 254 the unsubdividable
 * depth; any other depth addresses the boxes of that depth (none beyond the deepest leaf).
 *
 * Threading: one tree or matrix handle must not be used from two threads at once (B12 is not
 * thread-safe either, ImplicitBoxTree.h:116-122). Distinct handles are independent; each tree
 * owns a HIP stream. Calls return after their GPU work is complete.
 */
#ifndef B12X_C_API_H
#define B12X_C_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct b12x_tree_s* b12x_tree;
typedef struct b12x_scheme_s* b12x_scheme;
typedef struct b12x_coo_s* b12x_coo;
typedef struct b12x_krylov_s* b12x_krylov;

enum {
  B12X_OK = 0,
  B12X_E_ARG = -1,          /* invalid argument */
  B12X_E_HIP = -2,          /* HIP runtime error (message names the call) */
  B12X_E_NOMEM = -3,        /* device allocation failed */
  B12X_E_UNSUPPORTED = -4,  /* depth or map/grid combination outside this implementation */
  B12X_E_NODEVICE = -5,     /* no usable GPU */
  B12X_E_IO = -6            /* a file could not be opened, read or written (message says which) */
};

/* Flags.h:11-22 */
enum { B12X_NONE = 0x00, B12X_HIT = 0x01, B12X_INS = 0x02, B12X_EXPD = 0x04, B12X_SD = 0x08,
       B12X_EQU = 0x10, B12X_ENTR = 0x20, B12X_EXTR = 0x40, B12X_ALL = 0xFF };
#define B12X_LEAF_DEPTH 255 /* TypeDefinitions.hpp getLeafDepth */
#define B12X_INVALID_NRBOXES 0xFFFFFFFFu /* TypeDefinitions.hpp getInvalidNrBoxes */

/* Grid kinds (Grids.h:333-370); params are uint64:
 *   B12X_GRID_INNER   {nPointsPerDim}       InnerGrid(n), k = n^dim      (Grids.hpp:99-134)
 *   B12X_GRID_UNIFORM {nPoints, seed}       UniformGrid(k, seed), host (mt19937) semantics
 *                                           (Grids.hpp:136-176)                              */
enum { B12X_GRID_INNER = 0, B12X_GRID_UNIFORM = 1 };

/* Map kinds; params are double (missing trailing params take the defaults shown):
 *   B12X_MAP_IDENTITY       {}                              Map_Identity.h:9-21, any dim
 *   B12X_MAP_HENON          {a=1.4, b=0.3, steps=1}         Map_Henon.h:26-40, dim 2
 *   B12X_MAP_LORENZ         {sigma=10, rho=28, beta=8/3, h=0.01, steps=20}
 *                                                           Map_Lorenz.h:29-44 (RK4Auto), dim 3
 *   B12X_MAP_HENON_HEILES   {order=2, h=0.02, steps=50}     SymplecticIntegratorAuto<4>
 *                           (MapTools.hpp:361-523) with the Hénon–Heiles force; no
 *                           counterpart map in the reference, dim 4 (q1, q2, p1, p2)
 *   B12X_MAP_DUFFING        {delta=0.25, alpha=-1, beta=1, gamma=0.3, omega=1, t0=0,
 *                            tEnd=2*pi, h=2*pi/100}  RK4 (MapTools.hpp:35-108) of
 *                           x'' + delta x' + alpha x + beta x^3 = gamma cos(omega t); no
 *                           counterpart map in the reference, dim 2                           */
enum { B12X_MAP_IDENTITY = 0, B12X_MAP_HENON = 1, B12X_MAP_LORENZ = 2, B12X_MAP_HENON_HEILES = 3,
       B12X_MAP_DUFFING = 4 };

const char* b12x_last_error(void);
/* Name of the current device, for logs (ThrustSystem.hpp:52-62 printDevice). */
int b12x_device_name(char* buf, int len);

/* --- tree ------------------------------------------------------------------------------- */
/* ImplicitBoxTree(center, radius[, sdScheme]) — ImplicitBoxTree.h:49-56, .hpp:120-145.
 * sd_scheme may be NULL (default i % dim); otherwise n_sd entries repeated periodically. */
int b12x_tree_create(int dim, const double* center, const double* radius, const uint8_t* sd_scheme, int n_sd,
                     b12x_tree* out);
int b12x_tree_destroy(b12x_tree t);
/* The tree's dimension, centre and radius (dim doubles each) and subdivision scheme (sdScheme_:
 * *n_sd_out entries, 64 for the default scheme and 65 for a given one, as ImplicitBoxTree.hpp:
 * 27-33,3514-3545 size it; sd_scheme holds at least 65): the getters behind
 * ImplicitBoxTree::save (ImplicitBoxTree.hpp:460-486). Any output may be NULL. */
int b12x_tree_geometry(b12x_tree t, int* dim, double* center, double* radius, uint8_t* sd_scheme, int* n_sd_out);
/* setCenter / setRadius / setSdScheme — ImplicitBoxTree.h:550-560, .hpp:3464-3552. center and
 * radius (dim entries each) may be NULL to keep them; a negative radius is rejected (the tree is
 * unchanged). With set_sd != 0 the scheme becomes sd_scheme (n_sd entries repeated periodically,
 * or the default scheme when n_sd is 0); values >= dim are rejected. The leaves stay. */
int b12x_tree_set_domain(b12x_tree t, const double* center, const double* radius, const uint8_t* sd_scheme, int n_sd,
                         int set_sd);
/* areAllLeavesOnDepth(depth) — ImplicitBoxTree.h:84-85, .hpp:589-609 */
int b12x_all_leaves_on_depth(b12x_tree t, uint8_t depth, int* out);
/* a.hasSameBoxes(b) — ImplicitBoxTree.h:141-143, .hpp:1299-1381: same dimension, centre, radius,
 * leaves (a's leaves against b's first leaves, as the reference compares them) and scheme up to
 * a's depth(). */
int b12x_has_same_boxes(b12x_tree a, b12x_tree b, int* out);
/* getNHostBufferBytes / setNHostBufferBytes — ImplicitBoxTree.h:568-569: kept and written to the
 * tree's files (default 4 GiB). The reference consults it only on its host architecture
 * (Memory<HOST>::getFreeBytes, ThrustSystem.hpp:21-42); its device path sizes passes from the free
 * device memory (ThrustSystem.hpp:64-68), as this library does, so it limits nothing here. */
int b12x_get_host_buffer_bytes(b12x_tree t, uint64_t* out);
int b12x_set_host_buffer_bytes(b12x_tree t, uint64_t bytes);
/* Pass sizing of transitionMatrix (ImplicitBoxTree.hpp:2941-3004 sizes each pass from
 * Memory<DEVICE>::getFreeBytes): the test points of the source boxes are imaged and reduced in
 * passes over whole boxes, each holding at most `bytes` of per-point scratch (4 bytes a point for
 * k <= 64 test points per box, 24 above; at least one box a pass). 0, the default, bounds a pass
 * at a third of the device's free memory when it starts, as the reference sizes its first pass.
 * The matrix does not depend on the passes; its own arrays (16 bytes an entry, 12 more while the
 * column-major sort runs) are the memory that remains. No C ABI counterpart (the reference has
 * no setter: it always sizes from free memory). */
int b12x_set_tm_pass_bytes(b12x_tree t, uint64_t bytes);
int b12x_get_tm_pass_bytes(b12x_tree t, uint64_t* bytes);
/* Measurement (no counterpart in the reference): the passes of the calling thread's last
 * transitionMatrix. */
int b12x_tm_last_passes(uint32_t* passes);
/* size(), count(depth), depth() — ImplicitBoxTree.h:110-111,129-130,603 */
int b12x_size(b12x_tree t, uint32_t* out);
int b12x_count(b12x_tree t, uint8_t depth, uint32_t* out);
int b12x_depth(b12x_tree t, uint8_t* out);
/* The continuation algorithm of continuation.m:13-44 at one depth (no counterpart in the C++
 * reference; the MATLAB driver runs it through the MEX calls): deleteDepth(0), insertByPoints(points,
 * depth, INS), then while count(depth) grows {changeFlags(INS, EXPD, depth); insertByIterators(
 * ByFlags(depth, EXPD), depth, INS, NONE); unsetFlags(EXPD)}. counts receives count(depth) after
 * the first insertion and after each step (at most max_counts of them); *n_counts is their number. */
int b12x_continuation(b12x_tree t, b12x_scheme s, uint8_t depth, const double* points, uint64_t n_points,
                      uint32_t* counts, int max_counts, int* n_counts);
/* ImplicitBoxTree::boxes(Depth) (ImplicitBoxTree.h:86-91, ImplicitBoxTree.hpp:611-784): *n_out
   boxes of depth `depth`, each 2*dim+1 doubles {centre, radius, flags}, in leaf order; the root
   carries the AND of all leaf flags, a box between root and leaves the AND over its leaves.
   out may be NULL to ask for *n_out only. */
int b12x_boxes(b12x_tree t, uint8_t depth, double* out, uint64_t* n_out);
/* ImplicitBoxTree::enumerate() (ImplicitBoxTree.h:132-133, ImplicitBoxTree.hpp:1273-1283):
   *n_out = depth() + 1 entries count(0), ..., count(depth()); out may be NULL. */
int b12x_enumerate(b12x_tree t, uint32_t* out, int* n_out);
/* Copy the leaves out / in as B12's SoA (bitPatternVector_, depthVector_, flagsVector_,
 * ImplicitBoxTree.h:619-624). upload sorts the leaves into B12's order. */
int b12x_tree_download(b12x_tree t, uint64_t* bits, uint8_t* depth, uint8_t* flags);
int b12x_tree_upload(b12x_tree t, uint32_t n, const uint64_t* bits, const uint8_t* depth, const uint8_t* flags);

/* subdivide(flags) — ImplicitBoxTree.h:487-489 (returns #leaves with flags set) */
int b12x_subdivide(b12x_tree t, uint8_t flags, uint32_t* ret);
/* unsubdivide(flags) — ImplicitBoxTree.h:543-545 */
int b12x_unsubdivide(b12x_tree t, uint8_t flags, uint32_t* ret);
/* remove(flags) — ImplicitBoxTree.h:432-434 */
int b12x_remove(b12x_tree t, uint8_t flags, uint32_t* ret);
/* setFlags / unsetFlags / changeFlags(…, depth) — ImplicitBoxTree.h:461-463, 527-529, 93-94 */
int b12x_set_flags(b12x_tree t, uint8_t flags, uint8_t depth, uint32_t* ret);
int b12x_unset_flags(b12x_tree t, uint8_t flags, uint8_t depth, uint32_t* ret);
int b12x_change_flags(b12x_tree t, uint8_t flags_to_unset, uint8_t flags_to_set, uint8_t depth, uint32_t* ret);
/* *ByStencil with a host bool array of count(depth) bytes — ImplicitBoxTree.h:482-485, 540-541,
 * 106-108. op: 0 set, 1 unset, 2 change (flags2 used by change only). */
int b12x_flags_by_stencil(b12x_tree t, int op, const uint8_t* stencil, uint8_t flags, uint8_t flags2, uint8_t depth,
                          uint32_t* ret);

/* Energy-band predicate (no counterpart in the reference: the Hénon-Heiles configuration's
 * box predicate). op and flags as b12x_flags_by_stencil, with the stencil computed on the GPU:
 * leaf i is selected iff the interval bound of H(q,p) = (p1^2+p2^2)/2 + (q1^2+q2^2)/2 + q1^2 q2
 * - q2^3/3 over its box meets [emin, emax]. 4-dimensional trees, leaf depth only. */
int b12x_flags_by_energy_band(b12x_tree t, double emin, double emax, int op, uint8_t flags, uint8_t flags2,
                              uint8_t depth, uint32_t* ret);
/* *ByIterators over the image points of (scheme, depth_it, flags_to_check) —
 * setFlagsByIterators ImplicitBoxTree.h:465-469 (= setFlagsBySearch(searchCompacted(...))),
 * unsetFlagsByIterators :531-532, changeFlagsByIterators :96-98. */
int b12x_flags_by_image(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, int op, uint8_t flags,
                        uint8_t flags2, uint8_t depth, uint32_t* ret);
/* Number of points of the iterator pair (points.second - points.first). */
int b12x_num_points(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, uint64_t* out);
/* searchByIterators(points, depth, out) — ImplicitBoxTree.h:439-444; out has num_points entries */
int b12x_search_by_image(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, uint8_t depth,
                         uint32_t* out);
/* searchCompacted(points, depth) — ImplicitBoxTree.h:458-459: sorted unique hit indices.
 * out may be NULL to query n_out; otherwise it must hold count(depth) entries. */
int b12x_search_compacted(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, uint8_t depth,
                          uint32_t* out, uint64_t* n_out);
/* searchByPoints(points, depth) — ImplicitBoxTree.h:448-453; points row-major n x dim */
int b12x_search_by_points(b12x_tree t, const double* points, uint64_t n, uint8_t depth, uint32_t* out);

/* One subdivision step of subdivision.m:18-25 (subdivide(sd_flag); setFlagsByIterators(
 * OverLeaves, HIT); remove(HIT); unsetFlags(HIT)) fused into three kernels; steps are repeated
 * n_steps times. counts (may be NULL) receives 4 values per step: the return values of the
 * four reference calls. */
int b12x_subdivision_steps(b12x_tree t, b12x_scheme s, uint8_t sd_flag, int n_steps, uint32_t* counts);

/* Measurement hook (no counterpart in the reference; used by the benchmark): enable=1 clears
 * and enables HIP-event timing of every image-kernel launch on the tree's stream; the totals
 * since then are returned (milliseconds, launches, test points). enable=0 disables timing. */
int b12x_tree_image_timing(b12x_tree t, int enable, double* ms, uint64_t* launches, uint64_t* points);
/* Measurement (no counterpart in the reference): each image launch timed since timing was last
 * enabled, in launch order: its duration (ms) and test points. *n receives their number; at most
 * cap of them are written (ms or points may be NULL). */
int b12x_tree_image_launches(b12x_tree t, double* ms, uint64_t* points, uint64_t cap, uint64_t* n);

/* Measurement hook (no counterpart in the reference): `reps` launches of a plain device copy of
 * `bytes` bytes (16-byte vector loads and stores) between two fresh buffers; the average launch
 * time in milliseconds by HIP events. Used to calibrate HBM traffic counters. */
int b12x_bench_copy(uint64_t bytes, int reps, double* ms);

/* Measurement hook (no counterpart in the reference): a grid of 16384 x 256 lanes, each running
 * eight independent FP64 multiply-add chains, with the shader-clock and 100 MHz real-time
 * counters read around the loop. *ghz receives the shader clock under that full FP64 load; *frac
 * (may be NULL) the fraction of the SIMDs' FP64 issue slots at that clock the launch used. */
int b12x_bench_fp64_clock(double* ghz, double* frac);

/* --- scheme ----------------------------------------------------------------------------- */
int b12x_scheme_create(int dim, int grid_kind, const uint64_t* grid_params, int n_grid_params, int map_kind,
                       const double* map_params, int n_map_params, b12x_scheme* out);
int b12x_scheme_destroy(b12x_scheme s);
/* Grid::getNumberOfPointsPerBox — Grids.h:76 */
int b12x_scheme_points_per_box(b12x_scheme s, uint64_t* out);
/* Apply the map to n host points (row-major n x dim) on the GPU — the functor protocol of
 * doc/B12_doc.tex:109-146, exposed for integrator checks. */
int b12x_map_points(b12x_scheme s, const double* in, double* out, uint64_t n);

/* --- insertion --------------------------------------------------------------------------- */
/* ImplicitBoxTree::deleteDepth(depth) (ImplicitBoxTree.h:124-126, .hpp:1215-1262): every leaf
 * truncated to depth-1 with equal boxes merged (flags ANDed); depth 0 leaves the surrounding box
 * alone. *ret = count(depth) before (1 for depth 0; 0 for the leaf or unsubdividable depth or a
 * depth below every leaf). */
int b12x_delete_depth(b12x_tree t, uint8_t depth, uint32_t* ret);
/* ImplicitBoxTree::insertByIterators(points, depth, flagsForInserted, flagsForHit[, res])
 * (ImplicitBoxTree.h:170-187, .hpp:1613-1892) for the iterator of (depth_it, flags_to_check,
 * scheme), and insertByPoints (ImplicitBoxTree.h:189-202, .hpp:1894-1930) for n host points of
 * dim doubles each. res (may be NULL) receives isToInsert per point. */
int b12x_insert_by_image(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, uint8_t depth,
                         uint8_t flags_for_inserted, uint8_t flags_for_hit, uint8_t* res);
int b12x_insert_by_points(b12x_tree t, const double* points, uint64_t n, uint8_t depth, uint8_t flags_for_inserted,
                          uint8_t flags_for_hit, uint8_t* res);

/* --- files ---------------------------------------------------------------------------------- */
/* ImplicitBoxTree::save(fileName) (ImplicitBoxTree.h:70-72, .hpp:435-524): B12's JSON layout
 * (version 3.1, centre, radius, sdScheme, nLeaves, bitPatternVector, depthVector, flagsVector,
 * nHostBufferBytes); ".json" is appended when the name lacks it. */
int b12x_tree_save_json(b12x_tree t, const char* file_name);
/* ImplicitBoxTree::saveForGaio(fileName) (ImplicitBoxTree.h:74-76, .hpp:526-569): binary file
 * for GAIO's readForGaio.m (dim, sizeof(double), centre, radius, 500*dim+1 scheme entries as
 * int, then per leaf its centre, depth and flags). */
int b12x_tree_save_gaio(b12x_tree t, const char* file_name);
/* ImplicitBoxTree(fileName) (ImplicitBoxTree.h:58-59, .hpp:142-339): a JSON file (or the name
 * with ".json" appended when the name itself cannot be opened), or a legacy binary file of B12a
 * or B12b when the name does not end in ".json". Equal consecutive boxes are merged with their
 * flags ANDed, as the reference's loader does. dim is the tree's dimension (a template argument
 * in the reference). */
int b12x_tree_load(const char* file_name, int dim, b12x_tree* out);
/* CooMatrix::save(fileName) and CooMatrix(fileName) (CooMatrix.h:75-76,110-112,
 * CooMatrix.hpp:235-284,473-534): nRows, nColumns, nnz, rowIndices, columnIndices, values. */
int b12x_coo_save_json(b12x_coo m, const char* file_name);
int b12x_coo_load_json(const char* file_name, b12x_coo* out);

/* --- transition matrix -------------------------------------------------------------------- */
/* transitionMatrixForLeaves/ForRoot<double>(grid, map, useAbsoluteValues, useRowMajor,
 * useInvalidNrBoxes) — ImplicitBoxTree.h:491-525, .hpp:2932-3211. */
int b12x_transition_matrix(b12x_tree t, b12x_scheme s, uint8_t depth_it, uint8_t flags_to_check, uint8_t depth,
                           int use_abs, int row_major, int use_invalid, b12x_coo* out);
int b12x_coo_destroy(b12x_coo m);
/* nnz, nRows_, nColumns_ — CooMatrix.h:118-120 */
int b12x_coo_info(b12x_coo m, uint64_t* nnz, uint32_t* n_rows, uint32_t* n_cols, int* row_major);
/* rowIndices_, columnIndices_, values_ to host (any pointer may be NULL) */
int b12x_coo_download(b12x_coo m, uint32_t* rows, uint32_t* cols, double* vals);
/* Build a matrix from host COO arrays (CooMatrix.h:66-69). */
int b12x_coo_upload(const uint32_t* rows, const uint32_t* cols, const double* vals, uint64_t nnz, uint32_t n_rows,
                    uint32_t n_cols, b12x_coo* out);
/* CooMatrix::isSortedRowMajor / isSortedColumnMajor — CooMatrix.h:80-83, .hpp:354-374 (equal
 * neighbouring entries count as sorted) */
int b12x_coo_is_sorted(b12x_coo m, int row_major, int* out);
/* CooMatrix::sortRowMajor / sortColumnMajor — CooMatrix.h:88-91, .hpp:95-181: in place, stable */
int b12x_coo_sort(b12x_coo m, int row_major);
/* CooMatrix::makeSymmetric(makeAntisymmetric) — CooMatrix.h:84-86, .hpp:376-426: the symmetric
 * part (P + P^T)/2, or the antisymmetric part (P - P^T)/2, of a square matrix; equal entries are
 * summed from the last backwards; row-major order when P was sorted row-major, else
 * column-major */
int b12x_coo_make_symmetric(b12x_coo m, int antisymmetric);
/* CooMatrix copy constructor — CooMatrix.h:61-62: a device copy of the matrix */
int b12x_coo_copy(b12x_coo m, b12x_coo* out);
/* compressIndexVector(indices, dim) — CooMatrix.h:20-23: sorted index vector -> dim+1 offsets */
int b12x_compress_index_vector(const uint32_t* indices, uint64_t n, uint32_t dim, uint32_t* out);
/* Invariant measure by power iteration: mu_k = mu_0 P^k / ||mu_0 P^k||_1, mu_0 = 1/n (or the given
 * mu if mu_init != 0; mu holds n_rows doubles), until max_iter iterations or
 * ||mu_{k+1} - mu_k||_1 <= tol (tol < 0: fixed count). The iterates are kept unnormalised (scaled
 * by exact powers of two) and every sum runs in one fixed order, so the result does not depend on
 * the form that ran it. The reference computes this step in MATLAB (subdivisionByNim.m:30-59);
 * there is no C++ API. */
int b12x_power_iteration(b12x_coo m, int max_iter, double tol, int mu_init, double* mu, int* iters, double* resid);
/* b12x_power_iteration without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and kernels as b12x_power_iteration; the matrix must be
 * sorted column-major (transitionMatrix's default order, or after b12x_coo_sort(m, 0)) and is read
 * in place. Everything else lives in a file-scope device array that each GPU holds until the process
 * ends (about 22 MB): matrices of more than 131072 columns are rejected with B12X_E_UNSUPPORTED.
 * All calls on one GPU share that array, so calls that overlap on one GPU race; each call
 * initialises what it reads. mu holds n_rows doubles. */
int b12x_power_iteration_noalloc(b12x_coo m, int max_iter, double tol, int mu_init, double* mu, int* iters,
                                 double* resid);
/* Measurement (no counterpart in the reference): the duration of the calling thread's last
 * persistent power-iteration launch, from HIP events on the matrix's stream (0 when the last
 * call ran the batched form). */
int b12x_power_iteration_kernel_ms(double* ms);
/* Measurement (no counterpart in the reference): the form of the calling thread's last persistent
 * launch: 0 straight-line (one group of 64 columns per wave, columns of at most 24 entries), 1
 * waves sweeping several groups; -1 when the last call ran the batched form alone. */
int b12x_power_iteration_form(int* form);
/* Subspace iteration for the nvec (<= 8) leading eigenvalues (no counterpart in the reference,
 * which leaves eigenvalues to MATLAB eigs, subdivisionByNim.m:30-59): X_0 = cholQR2([1/n, splitmix64
 * columns]); for k = 0, 1, ...: Y = X_k P (row-vector convention, as the power iteration), the Ritz
 * matrix H = X_k^T Y and the residual of the leading nlead (1 <= nlead <= nvec) columns,
 * rho = ||Y_L - X_L H_LL||_F / ||H_LL||_F (L: those columns; H_LL the leading nlead x nlead block),
 * zero when their span is invariant (Cholesky QR keeps the spans of leading columns nested, so
 * they converge to the dominant invariant subspace of that dimension); it stops at the first k
 * with rho <= tol (tol >= 0) or k = max_iter, else X_{k+1} = cholQR2(Y). ritz (nvec x nvec,
 * row-major) receives that H, whose eigenvalues approximate the leading eigenvalues of P; basis
 * (nvec x n, column by column, may be NULL) X_k; *iters (may be NULL) k, the basis updates made;
 * *resid (may be NULL) rho. Every sum runs in one fixed order, so results do not depend on
 * timing. */
int b12x_subspace_iteration(b12x_coo m, int nvec, int nlead, int max_iter, double tol, double* ritz, double* basis,
                            int* iters, double* resid);
/* Krylov basis for a Krylov-Schur eigensolver (no counterpart in the reference, which leaves its
 * eigenvalues to MATLAB eigs, subdivisionByNim.m:30-56; a host driver, eigs() in b12_amd/api.py,
 * does the small dense steps). The operator is the row-vector product w = v P over the matrix's
 * columns < n_rows (w_j = sum over column j in ascending row order, as the power iteration).
 * Basis V: ncv + 1 (3 <= ncv + 1 <= 64) device vectors of n_rows doubles; the matrix may be
 * destroyed after create. Every dot product over the boxes runs in one fixed order (block-strided
 * partial sums, a fixed reduction tree), so results do not depend on timing.
 *  - start: V[:, 0] = v0 / ||v0|| (v0: host, n_rows doubles; NULL: v0_i = (splitmix64(8 i + 1)
 *    >> 11) 2^-53 * 2 - 1); *norm (may be NULL) ||v0||.
 *  - expand: w = V[:, p] P, then classical Gram-Schmidt against V[:, 0..p] twice (h1 = V^T w,
 *    w -= V h1 with one subtraction per column in column order; h2 likewise); h[c] = h1[c] + h2[c]
 *    (p + 1 values), *beta = ||w||, and V[:, p + 1] = w / ||w|| unless ||w|| = 0. 0 <= p < ncv.
 *  - rotate: V[:, c] = sum over d < m of V[:, d] q[d * k + c] (d ascending) for c < k, then
 *    V[:, k] = V[:, m]. q: host, m x k row-major; 1 <= k <= m <= ncv.
 *  - vector: V[:, c] to the host (n_rows doubles), 0 <= c <= ncv. */
int b12x_krylov_create(b12x_coo m, int ncv, b12x_krylov* out);
int b12x_krylov_destroy(b12x_krylov k);
int b12x_krylov_start(b12x_krylov k, const double* v0, double* norm);
int b12x_krylov_expand(b12x_krylov k, int p, double* h, double* beta);
int b12x_krylov_rotate(b12x_krylov k, int m, int kk, const double* q);
int b12x_krylov_vector(b12x_krylov k, int c, double* out);
/* Strong components of the matrix's graph (an edge row -> column per entry; entries naming a
 * vertex >= n_rows, such as the column of points outside the cover, are left out), as
 * chainRecurrentSet.m:26-39 obtains them from MATLAB's graphconncomp. labels (n_rows entries,
 * may be NULL): the largest vertex index of the vertex's component. n_components and
 * n_nontrivial (components of more than one vertex) may be NULL. */
int b12x_coo_scc(b12x_coo m, uint32_t* labels, uint32_t* n_components, uint32_t* n_nontrivial);
/* Measurement (no counterpart in the reference): of the calling thread's last strong-components
 * run (b12x_coo_scc, or the last step of b12x_chain_recurrent_steps), the colouring rounds, the
 * edge and vertex passes launched and the host reads of the change counters. */
int b12x_scc_last_stats(uint32_t* rounds, uint32_t* passes, uint32_t* checks);
/* The loop body of chainRecurrentSet.m:18-44, n_steps times, on the device: subdivide(sd_flag);
 * the leaves' transition matrix (transitionMatrix({grid, map, -1})); strong components; HIT on
 * the leaves whose component has more than one leaf; remove(HIT); unsetFlags(HIT). counts (may
 * be NULL) receives 4 values per step: the leaves after subdividing, the components, the
 * components of more than one leaf, the leaves kept. */
int b12x_chain_recurrent_steps(b12x_tree t, b12x_scheme s, uint8_t sd_flag, int n_steps, uint32_t* counts);

/* --- batched and served forms (configs 4 and 5) --------------------------------------------
 * A forest runs the subdivision algorithm (subdivision.m:15-25) for many jobs at once: jobs of
 * one (grid, map kind) share every launch, each job with its own domain, map parameters and
 * target depth. A job whose depth is reached retires: its transition matrix
 * (transitionMatrixForLeaves, ImplicitBoxTree.hpp:3178-3193; invalid targets dropped) and its
 * invariant measure (power iteration with pi_max_iter / pi_tol as in b12x_power_iteration) are
 * computed for all jobs retiring in the same step in shared launches. Results do not depend on
 * which jobs share a launch. These entry points have no counterpart in the reference (B12 runs
 * one tree per call). */
typedef struct b12x_forest_s* b12x_forest;
int b12x_forest_create(int dim, int grid_kind, const uint64_t* grid_params, int n_grid_params, int map_kind,
                       int use_abs, int row_major, int pi_max_iter, double pi_tol, b12x_forest* out);
int b12x_forest_destroy(b12x_forest f);
/* queue a job (root box flagged SD); it joins the next step */
int b12x_forest_add(b12x_forest f, const double* center, const double* radius, const double* map_params,
                    int n_map_params, uint8_t target_depth, uint32_t* job);
/* one merged subdivision step for every active job, then retirement of finished jobs */
int b12x_forest_step(b12x_forest f, uint32_t* n_active, uint32_t* n_retired);
int b12x_forest_run(b12x_forest f);
/* state: 0 queued, 1 subdividing, 4 iterating its invariant measure, 2 done, 3 released */
int b12x_forest_job_info(b12x_forest f, uint32_t job, int* state, uint32_t* n_leaves, uint64_t* nnz, int* iters,
                         double* resid);
int b12x_forest_job_cover(b12x_forest f, uint32_t job, uint64_t* bits, uint8_t* depth, uint8_t* flags);
int b12x_forest_job_matrix(b12x_forest f, uint32_t job, uint32_t* rows, uint32_t* cols, double* vals);
int b12x_forest_job_measure(b12x_forest f, uint32_t job, double* mu);
int b12x_forest_job_release(b12x_forest f, uint32_t job);
/* HBM bytes held by the forest's leaf arrays, tables, measure pool and unreleased results */
int b12x_forest_footprint(b12x_forest f, uint64_t* bytes);
/* Telemetry, first n of: steps, seconds subdividing, seconds retiring (transition matrices),
 * seconds iterating measures, measure iterations, the pool's peak jobs, columns, entries, and
 * the pool's columns and entries summed over the iterations launched (work of the measures). */
int b12x_forest_stats(b12x_forest f, double* out, int n);
/* Measure iterations the pool advances per forest step (default 64). */
int b12x_forest_set_pool_iterations(b12x_forest f, int iters);
/* Continuous-batching request loop over n_forests forests: request i arrives arrival[i] seconds
 * after the start (sorted), goes to forests[forest_of[i]] with centre/radius (stride doubles
 * apart), map_params (n_map_params each, may be NULL) and depth. It is admitted once it has
 * arrived and the forests' footprint now (b12x_forest_footprint: arrays, pool and unreleased
 * results) plus the projected peak of every job in flight and of itself fits budget_bytes (a job
 * alone is always admitted). A job's projected peak is the cost of the largest finished job of
 * its forest and depth (leaves x (208 + 4k) bytes + matrix entries x 74 bytes: leaf buffers with
 * their growth slack, pool, result arena and the transition matrix's transients), else of the
 * deepest smaller depth doubled per level, else reserve_per_job. Every pass runs one merged step
 * of each forest with work. latency[i] = completion[i] - arrival[i] (seconds); job_ids (may be
 * NULL) gets each request's job id in its forest, admitted (may be NULL) its admission time;
 * stats (may be NULL) receives the peak footprint seen (bytes), the most jobs in flight at once
 * and how many times a waiting request was held back by the budget. */
int b12x_serve_run(b12x_forest* forests, int n_forests, uint32_t n_req, const double* arrival, const int* forest_of,
                   const double* centers, const double* radii, int stride, const double* map_params,
                   int n_map_params, const uint8_t* depths, double budget_bytes, double reserve_per_job,
                   double* latency, double* completion, uint32_t* job_ids, double* admitted, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* B12X_C_API_H */
