/*
 * kmeans_amd.h -- C-ABI of the MI355X (gfx950) K-means engine.
 *
 * Three layers, all extern "C", plain pointers and sizes:
 *
 *  1. Host API, the drop-in for the reference's public entry points
 *     (main/cpp/kmeans.hpp:44-132): computeKmeansF / computeKmeansD with the
 *     same argument lists and status codes, kmeansGetErrorString, plus the
 *     handle-based engine (kmeansCreate ... kmeansDestroy) that exposes what
 *     the reference computes but never returns (labels, compactness, initial
 *     centres supplied by the caller, per-phase timings).
 *
 *  2. Kernel API (km*), the replacement for the reference's kernel C-ABI
 *     (main/cuda/KmeansCudaKernels.h:170-478): device pointers, an explicit
 *     HIP stream (passed as void*, 0 = null stream), caller-sized workspace
 *     with size queries, no process-global state, reentrant.  Nothing here
 *     allocates device memory.
 *
 *  3. Legacy kernel names (ClosestCentersF, ClusterCentersF, CompactnessF,
 *     rowNormalizeF, colNormalizeF and D variants) with the reference's
 *     argument lists, so existing callers link unchanged.
 *
 * Layouts: points X are N x D row-major (the reference's "m x n C-ordered"
 * data, main/cpp/kmeans.hpp:88-92).  Engine centres are K x D row-major; the
 * legacy API's result buffer is D x K (main/cpp/kmeans.cpp:886-888).
 *
 * Arithmetic contract (the plain FP32 path): ||c_j||^2 and the score
 * s_ij = -0.5f*||c_j||^2 + x_i.c_j are single-rounding fma chains over d in
 * ascending order; the label is the lowest j among the maxima of s_ij.  New
 * centres and the compactness are FP64 sums in a fixed order (KM_CHUNK-member
 * pieces within KM_BLOCK-point blocks), so results do not depend on the GPU,
 * the launch shape or the N-tiling.
 */
#ifndef KMEANS_AMD_H_
#define KMEANS_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KM_CHUNK 256           /* members per first-level FP64 partial sum */
#define KM_BLOCK (1L << 23)    /* points per canonical summation block     */

/* C++ callers get the reference's default arguments (kmeans.hpp:102-107,
 * 125-130); C callers pass every argument. */
#ifdef __cplusplus
#define KM_DEFAULT_ARG(v) = v
#else
#define KM_DEFAULT_ARG(v)
#endif

/* ---- status codes: values of main/cpp/kmeans.hpp:44-59 ------------------ */
typedef enum kmeansErrorStatus {
  KMEANS_SUCCESS = 0,
  KMEANS_ERROR_BUILDDATA = 1,
  KMEANS_ERROR_INITIALIZE = 2,
  KMEANS_ERROR_CLOSEST_CENTERS = 3,
  KMEANS_ERROR_CLUSTER_CENTERS = 4,
  KMEANS_ERROR_COMPACTNESS = 5,
  KMEANS_ERROR_TRANSPOSE = 6,
  KMEANS_ERROR_RESET = 7,
  KMEANS_ERROR_COMPUTE = 8,
  KMEANS_ERROR_GET_CENTERS = 9,
  KMEANS_ERROR_TILING = 10,
  KMEANS_ERROR_COPY_TILE = 11,
  KMEANS_ERROR_CONSTRUCT_MINIBATCH = 12,
  KMEANS_NONCONVERGENCE = 13,
  KMEANS_ERROR_INVALID_ARGUMENT = 14 /* new: shape or option rejected */
} kmeansErrorStatus;

/* ---- kernel status codes: values of KmeansCudaKernels.h:174-183 --------- */
typedef enum kmeansCudaErrorStatus {
  NO_ERROR = 0,
  ERROR_MATMATMULT = 1,
  ERROR_ROWTRANSFORMMINIMUM = 2,
  ERROR_CLOSESTCENTERS = 3,
  ERROR_CLUSTERCENTERS = 4,
  ERROR_COMPACTNESS = 5,
  ERROR_ROWNORMALIZE = 6,
  ERROR_COLNORMALIZE = 7,
  ERROR_INVALID_ARGUMENT = 8, /* new: bad shape, null pointer, small workspace */
  ERROR_UNSUPPORTED = 9       /* new: no GPU / not built for this device     */
} kmeansCudaErrorStatus;

/* =========================================================================
 * 1. Host API
 * ========================================================================= */

/* Replaces kmeansGetErrorString (main/cpp/kmeans.hpp:66-82). */
const char *kmeansGetErrorString(kmeansErrorStatus err);

/* Replaces computeKmeansF (main/cpp/kmeans.hpp:102-107, kmeans.cpp:1316-1361).
 * data: m x n host, C-ordered.  result: n x k host (D x K).
 * initAlgorithm: the reference ignores it and always starts from the rows
 * its own shuffle and srand(42)/rand() pick select (kmeans.cpp:553-556,
 * 876-888); here 0 and every value other than 1, 2 and 3 start from those
 * same rows (kmeansReferenceInitRows), 1 takes the initial centres from
 * result (D x K), 2 seeds them by k-means++ and 3 samples k distinct rows
 * (Floyd's algorithm, seed 42).
 * Deliberate deltas: the rows are not reordered (labels come back in the
 * caller's order and the canonical sums run over the caller's order); an
 * empty cluster keeps its previous centre instead of becoming NaN.
 * doFullSGEMM == 1 assigns through a library GEMM (rocBLAS) and a row argmin
 * as kmeans.cpp:918-992 does; otherwise the fused kernels.  matrixFormat and
 * numRetries are ignored (the reference ignores numRetries too).
 * miniBatchFraction < 1 runs the mini-batch loop (kmeans.cpp:1253-1312). */
kmeansErrorStatus computeKmeansF(const float *data, int m, int n, int k, float criterion,
                                 int maxIters, int numRetries, int initAlgorithm,
                                 int doFullSGEMM, int matrixFormat, float *result,
                                 float miniBatchFraction KM_DEFAULT_ARG(1.0f));

/* Replaces computeKmeansD (main/cpp/kmeans.hpp:125-130, kmeans.cpp:1364-1409).
 * Same contract and deltas in double precision: scores are fma chains in
 * double, centres the FP64 canonical sums divided by the counts (no rounding
 * to float).  The reference's fused FP64 path gives label 0 to every row
 * (reduction.hcu:8-11); this one computes real labels. */
kmeansErrorStatus computeKmeansD(const double *data, int m, int n, int k, double criterion,
                                 int maxIters, int numRetries, int initAlgorithm,
                                 int doFullSGEMM, int matrixFormat, double *result,
                                 double miniBatchFraction KM_DEFAULT_ARG(1.0));

/* Assignment precision of the engine. */
typedef enum kmeansPrecision {
  KM_PREC_FP32 = 0,      /* plain path: the canonical fma-chain scores (run on
                            the FP32 matrix cores, bit-identical)            */
  KM_PREC_BF16 = 1,      /* BF16 MFMA scores + FP32 re-resolve of near-ties
                            (n <= 256)                                        */
  KM_PREC_FP16 = 2       /* FP16 MFMA scores + FP32 re-resolve of near-ties
                            (n <= 256)                                        */
} kmeansPrecision;

/* Candidate proposer of the assignment (no counterpart in the reference).
 * The centres are split once into `groups` groups by a coarse K-means of the
 * centres; after `train_iters` plain iterations each point is scored exactly
 * only against the members of `top_groups` proposed groups, its label is
 * certified by a triangle-inequality bound (centre separation, or the group
 * bound ||x - c|| >= ||x - G_g|| - r_g), and uncertified points are searched
 * in full, so labels equal the plain FP32 labels.  Resident data only (no
 * streamed tiles, no mini-batch). */
typedef enum kmeansProposer {
  KM_PROPOSER_NONE = 0,
  KM_PROPOSER_MLP = 1,       /* MLP n -> hidden -> hidden -> groups, trained on
                                group(label) of rows sampled in the plain
                                iterations (BF16 MFMA, FP32 masters, Adam)  */
  KM_PROPOSER_CENTROID = 2   /* the top_groups nearest group centroids      */
} kmeansProposer;

typedef struct kmeansOptions {
  int precision;          /* kmeansPrecision                                  */
  int device;             /* HIP device ordinal                               */
  void *stream;           /* hipStream_t to run on; NULL: engine creates one  */
  long tile_rows;         /* points per streamed tile (0: all resident)       */
  int timing;             /* 1: record hipEvents per phase (no host syncs)    */
  int proposer;           /* kmeansProposer                                   */
  int groups;             /* centre groups (0: min(128, k); MLP: multiple of 16) */
  int top_groups;         /* groups proposed per point (0: 4; at most 8)      */
  int train_iters;        /* plain iterations before the proposer (0: 3)      */
  int hidden;             /* MLP hidden width (0: 256; multiple of 16)        */
  int full_gemm;          /* 1: assign by a library GEMM (rocBLAS SGEMM) into
                             an HBM product and a row argmin, the reference's
                             doFullSGEMM = 1 path (FP32 precision only; labels
                             follow the library's rounding)                   */
} kmeansOptions;

typedef struct kmeansStats {
  int iterations;
  double compactness;     /* FP64, against the final centres                  */
  double rel_err;
  long long empty_clusters;   /* summed over iterations                       */
  long long reresolved;       /* points re-searched in FP32 (MFMA paths)      */
  long long rescored;         /* points certified by exact candidate scores   */
  double ms_assign;           /* summed over iterations (timing == 1)         */
  double ms_reresolve;
  double ms_update;
  double ms_compactness;
  double ms_total;
  long long proposed;         /* point assignments made through the proposer  */
  long long certified;        /* ... certified by the centre-separation bound */
  long long certified_group;  /* ... certified by the group bound             */
  long long fallback;         /* ... searched in full (FP32)                  */
  long long candidate_pairs;  /* exact (point, centre) scores in named groups */
  long long proposal_hits;    /* ... whose final label was in a named group   */
  double ms_train;            /* grouping + MLP training (once, wall clock)   */
  double train_loss_first;    /* MLP cross-entropy, first and last batch      */
  double train_loss_last;
  long long deep;             /* points given to the split-BF16 second pass   */
  long long tiles;            /* (256-point, 32-centre) tiles the BF16/FP16
                                 assign computed; the rest were cut by the
                                 triangle bound                               */
  long long tiles_total;      /* ... out of this many                         */
  long long bounded;          /* point assignments kept by the drift bounds   */
  long long paired;           /* ... decided between the label and the runner-up
                                 by their two FP32 chains, the drift bounds
                                 having separated every other centre          */
  long long scored;           /* points the BF16/FP16 assign kernel scored
                                 against every centre                         */
  long long assign_launches;  /* launches of that kernel (timing == 1)        */
  double ms_assign_kernel;    /* their summed event-timed durations           */
  long long graph_launches;   /* iterations run as one hipGraph launch (FP32,
                                 resident single-block data, timing == 0)     */
} kmeansStats;

typedef struct kmeansEngine *kmeansHandle;

/* Engine lifecycle.  m points of dimension n, k centres: 1 <= k <= m,
 * m < 2^31, k <= 16384; the BF16/FP16 precisions need n <= 256. */
kmeansErrorStatus kmeansCreate(kmeansHandle *h, long m, int n, int k, const kmeansOptions *opt);
void kmeansDestroy(kmeansHandle h);
/* Points: host (on_device == 0, copied) or device (borrowed, must outlive h). */
kmeansErrorStatus kmeansSetData(kmeansHandle h, const float *data, int on_device);
/* Initial centres, k x n host (C-ordered). */
kmeansErrorStatus kmeansSetCenters(kmeansHandle h, const float *centers);
/* Initial centres by k-means++ seeding (Arthur & Vassilvitskii 2007) on the
 * GPU, resident data only.  No counterpart in the reference (its
 * initAlgorithm is ignored, kmeans.cpp:876-888).  Deterministic for a seed:
 * distances in FP64, the draw walks the canonical 256-way sum tree. */
kmeansErrorStatus kmeansInitPlusPlus(kmeansHandle h, uint64_t seed);
/* The k rows the reference's initialize takes as initial centres, as indices
 * into the caller's (unshuffled) rows: rows[i] = indices[points[i]], where
 * indices is 0..m-1 shuffled by std::shuffle with std::default_random_engine(0)
 * (kmeans.cpp:553-556) and points[i] = rand() % m after srand(42), redrawn
 * while std::binary_search over the earlier points (an unsorted vector, as in
 * the reference) finds it (kmeans.cpp:876-883).  Computed with this
 * process's libstdc++ and glibc routines (rand's sequence from a private
 * state: the caller's rand() is untouched).  Host only, no device. */
kmeansErrorStatus kmeansReferenceInitRows(long m, int k, int *rows);
/* Lloyd loop: while (it < maxIters && relErr > criterion) -- kmeans.cpp:1194. */
kmeansErrorStatus kmeansRun(kmeansHandle h, int maxIters, double criterion);
kmeansErrorStatus kmeansGetCenters(kmeansHandle h, float *centers /* k x n host */);
kmeansErrorStatus kmeansGetLabels(kmeansHandle h, int *labels /* m host */);
kmeansErrorStatus kmeansGetStats(kmeansHandle h, kmeansStats *stats);
/* The reference class's steps, one call each (used by include/kmeans_amd.hpp):
 * initialize (kmeans.cpp:822-907): initAlgorithm 1 takes centers_dxk (n x k,
 * the reference's result layout), 2 seeds by k-means++, 3 samples k distinct
 * rows with the seed, any other value takes kmeansReferenceInitRows' rows;
 * closestCenters (:911-1014): one assignment of every resident point;
 * clusterCenters (:1019-1041): new centres from the current labels;
 * compactness (:1045-1104): FP64 sum of squared distances to the current
 * centres of the current labels.  Resident data only; each call returns when
 * its work is done. */
kmeansErrorStatus kmeansInitialize(kmeansHandle h, int initAlgorithm, uint64_t seed,
                                   const float *centers_dxk);
kmeansErrorStatus kmeansAssign(kmeansHandle h);
kmeansErrorStatus kmeansUpdateCenters(kmeansHandle h);
kmeansErrorStatus kmeansCompactness(kmeansHandle h, double *out);
/* Mini-batch mode (kmeans.cpp:1253-1312): each iteration assigns and updates
 * from the next window of `rows` points (wrapping at m); 0 restores the
 * standard loop over all points. */
kmeansErrorStatus kmeansSetMiniBatch(kmeansHandle h, long rows);
/* Device pointers owned by the engine (valid until kmeansDestroy). */
const float *kmeansDeviceCenters(kmeansHandle h);
const int *kmeansDeviceLabels(kmeansHandle h);
kmeansErrorStatus kmeansSynchronize(kmeansHandle h);

/* Double-precision engine handle: the same steps as the float handle for
 * the reference's kmeans<double> class (kmeans.hpp:182-312, instantiated in
 * kmeans.cpp:1364-1409).  Scores are fma chains in double, centres the FP64
 * canonical sums divided by the counts.  Options: device and full_gemm are
 * used; precision must be KM_PREC_FP32 (meaning: no reduced precision); no
 * streamed tiles and no proposer.  Stats fill iterations, compactness and
 * rel_err. */
typedef struct kmeansEngineD *kmeansHandleD;
kmeansErrorStatus kmeansCreateD(kmeansHandleD *h, long m, int n, int k, const kmeansOptions *opt);
void kmeansDestroyD(kmeansHandleD h);
kmeansErrorStatus kmeansSetDataD(kmeansHandleD h, const double *data, int on_device);
kmeansErrorStatus kmeansSetCentersD(kmeansHandleD h, const double *centers /* k x n host */);
kmeansErrorStatus kmeansInitializeD(kmeansHandleD h, int initAlgorithm, uint64_t seed,
                                    const double *centers_dxk);
kmeansErrorStatus kmeansAssignD(kmeansHandleD h);
kmeansErrorStatus kmeansUpdateCentersD(kmeansHandleD h);
kmeansErrorStatus kmeansCompactnessD(kmeansHandleD h, double *out);
kmeansErrorStatus kmeansRunD(kmeansHandleD h, int maxIters, double criterion);
kmeansErrorStatus kmeansSetMiniBatchD(kmeansHandleD h, long rows);
kmeansErrorStatus kmeansGetCentersD(kmeansHandleD h, double *centers /* k x n host */);
kmeansErrorStatus kmeansGetLabelsD(kmeansHandleD h, int *labels /* m host */);
kmeansErrorStatus kmeansGetStatsD(kmeansHandleD h, kmeansStats *stats);

/* =========================================================================
 * 1b. Batched independent clusterings (no counterpart in the reference: its
 * host class runs one problem per object).  For callers with thousands of
 * small problems of uneven size: every instance has its own points (n[i] rows,
 * concatenated in instance order), its own k centres and the same canonical
 * semantics as kmeansRun (FP32 labels, canonical FP64 sums, compactness from
 * FP64 moments).  Each phase of an iteration is one persistent launch over
 * all instances, and a run of `iters` iterations is one hipGraph launch.
 * Limits: d <= 64, k <= 1024, k * (d rounded up to 4, 8, 16, 32 or 64, + 1)
 * * 4 bytes <= 64 KiB, k <= n[i] < KM_BLOCK, sum of n[i] < 2^31.
 * Precision: KM_PREC_FP32, or KM_PREC_BF16 for d <= 32 (BF16 MFMA scores,
 * each label certified against the FP32 scores or searched exactly, so the
 * labels are the FP32 labels).
 * ========================================================================= */
typedef struct kmeansBatchStats {
  int iterations;          /* total over runs                                 */
  int last_run_iters;
  int tile_rows;           /* points per assign work item (autotuned)         */
  int work_items;
  int graph_nodes;
  int tuned_from_cache;    /* 1: tile_rows came from the tuning cache         */
  long long empty_clusters;
  double ms_graph;         /* device time of the last run's graph             */
  double ms_capture;       /* host time to capture and instantiate it         */
  float tune_ms[5];        /* assign time per candidate tile: 256 .. 4096     */
  long long exact_searched; /* BF16: point assignments certified by no bound
                               and searched exactly (summed over iterations) */
  long long kept;          /* BF16: point assignments kept by the drift bounds */
  long long paired;        /* BF16: ... decided between the label and the
                              runner-up by their two FP32 chains             */
  long long scored;        /* BF16: ... scored against every centre          */
  double ms_assign;        /* options.timing == 1 (the run is then enqueued
                              eagerly, not as a graph): the last run's assign
                              launches, summed over iterations (HIP events)  */
  double ms_update;        /* ... and its update launches                     */
} kmeansBatchStats;

typedef struct kmeansBatchEngine *kmeansBatchHandle;

kmeansErrorStatus kmeansBatchCreate(kmeansBatchHandle *h, int instances, const long *n, int d,
                                    int k, const kmeansOptions *opt);
void kmeansBatchDestroy(kmeansBatchHandle h);
/* All points (sum n[i] x d), host (copied) or device (borrowed; 16-byte
 * aligned when d is a multiple of 4, else KMEANS_ERROR_INVALID_ARGUMENT). */
kmeansErrorStatus kmeansBatchSetData(kmeansBatchHandle h, const float *X, int on_device);
/* Initial centres, instances x k x d, host or device. */
kmeansErrorStatus kmeansBatchSetCenters(kmeansBatchHandle h, const float *C, int on_device);
/* Initial centres = the first k points of each instance. */
kmeansErrorStatus kmeansBatchInitFirstK(kmeansBatchHandle h);
/* Points per assign work item (even); 0 = autotune on the next run. */
kmeansErrorStatus kmeansBatchSetTileRows(kmeansBatchHandle h, int rows);
/* `iters` Lloyd iterations of every instance (criterion < 0 semantics). */
kmeansErrorStatus kmeansBatchRun(kmeansBatchHandle h, int iters);
kmeansErrorStatus kmeansBatchGetCenters(kmeansBatchHandle h, float *C /* instances x k x d */);
kmeansErrorStatus kmeansBatchGetLabels(kmeansBatchHandle h, int *labels /* sum n[i] */);
/* Compactness of every instance after each iteration of the last run
 * (last_run_iters x instances, FP64). */
kmeansErrorStatus kmeansBatchGetCompactness(kmeansBatchHandle h, double *comp);
kmeansErrorStatus kmeansBatchGetStats(kmeansBatchHandle h, kmeansBatchStats *stats);
const float *kmeansBatchDeviceCenters(kmeansBatchHandle h);
const int *kmeansBatchDeviceLabels(kmeansBatchHandle h);

/* =========================================================================
 * 2. Kernel API (device pointers; stream is a hipStream_t or NULL)
 * ========================================================================= */

/* Replaces colNormalizeF (KmeansCudaKernels.h:458-463, Normalize.cu:33-47):
 * nc[j] = fma chain over d of C[j][d]^2.  C is k x d row-major. */
kmeansCudaErrorStatus kmColNorms(void *stream, int k, int d, const float *C, float *nc);

/* Padded leading dimension of the D x K centre copy the assign kernels read. */
int kmCentersLd(int k);

/* Transposes C (k x d) into Ct (d x kmCentersLd(k), zero padded). */
kmeansCudaErrorStatus kmTransposeCenters(void *stream, int k, int d, const float *C, float *Ct);

/* Replaces ClosestCentersF (KmeansCudaKernels.h:238-244, ClosestCenters.cu:
 * 46-123): fused score + argmax, no N x K or N x ceil(K/64) buffers.
 * rows == NULL: points 0..m-1 of X, labels[i] written for i < m.
 * rows != NULL: points rows[0..m-1] (m read from *m_dev when m_dev != NULL),
 * labels[rows[p]] written.  best (nullable) receives the winning score. */
kmeansCudaErrorStatus kmAssignF32(void *stream, long m, const int *m_dev, int d, const float *X,
                                  const int *rows, int k, const float *Ct, const float *nc,
                                  int *labels, float *best);

/* Workspace bytes for kmUpdate / kmCompactness at these sizes. */
size_t kmUpdateWorkspaceSize(long n, int d, int k);

/* Replaces ClusterCentersF (KmeansCudaKernels.h:337-343, ClusterCenters.cu:
 * 3-132): counts and canonical FP64 means.  C_new may alias C_old.  Ct_new
 * (nullable) receives the D x ld copy, nc_new (nullable) the new norms,
 * counts (nullable, k int64) the member counts, empty (nullable, device int)
 * the number of empty clusters (which keep their previous centre). */
kmeansCudaErrorStatus kmUpdate(void *stream, long n, int d, int k, const float *X,
                               const int *labels, const float *C_old, float *C_new,
                               float *Ct_new, float *nc_new, long long *counts, int *empty,
                               void *workspace, size_t workspace_bytes);

/* kmUpdate without a workspace argument.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The workspace
 * is a fixed device array of KM_NOALLOC_SCRATCH_BYTES inside the library
 * (one per device, held until the process ends), so shapes whose
 * kmUpdateWorkspaceSize(n, d, k) exceeds it are rejected with
 * ERROR_INVALID_ARGUMENT (about n = 1.5M points at d = 128, k = 1024).  The
 * call allocates nothing and clears its accumulators with a kernel.  All
 * calls on one device share the array: calls that overlap on one device
 * (other streams, other threads) race; calls on one stream are ordered.
 * Results are those of kmUpdate bit for bit. */
#define KM_NOALLOC_SCRATCH_BYTES (32L << 20)
kmeansCudaErrorStatus kmUpdateNoAlloc(void *stream, long n, int d, int k, const float *X,
                                      const int *labels, const float *C_old, float *C_new,
                                      float *Ct_new, float *nc_new, long long *counts,
                                      int *empty);

/* Replaces CompactnessF (KmeansCudaKernels.h:389-394, Compactness.cu:3-96):
 * out[0] (device double) = canonical FP64 sum of ||x_i - C[l_i]||^2. */
kmeansCudaErrorStatus kmCompactness(void *stream, long n, int d, int k, const float *X,
                                    const int *labels, const float *C, double *out,
                                    void *workspace, size_t workspace_bytes);

/* Replaces rowNormalizeF (KmeansCudaKernels.h:428-432, Normalize.cu:3-31):
 * out[m0+i] = fma chain over d of X[i][d]^2 for i < m. */
kmeansCudaErrorStatus kmRowNorms(void *stream, long m0, long m, int d, const float *X, float *out);

/* Synthetic Gaussian mixture rows [i0, i0+n) into X (device), counter-based
 * so any slice can be made independently; the same streams as the oracle's
 * generator.  means (ncomp x d) and sigmas (ncomp) are device arrays; cdf
 * (ncomp doubles, device) holds the cumulative component weights. */
kmeansCudaErrorStatus kmGenerateMixture(void *stream, uint64_t seed, long i0, long n, int d,
                                        int ncomp, const float *means, const float *sigmas,
                                        const double *cdf, float *X, int *z);

/* Gather rows: out[r] = X[idx[r]] for r < n (device). */
kmeansCudaErrorStatus kmGatherRows(void *stream, long n, int d, const float *X, const int *idx,
                                   float *out);

/* Candidate-group proposer MLP (din -> hidden -> hidden -> groups, ReLU,
 * softmax cross-entropy; BF16 products, FP32 accumulation and parameters).
 * No counterpart in the reference: it serves the learned pruning of the
 * assignment.  Parameters are one flat FP32 vector: W1 (hidden x din), b1,
 * W2 (hidden x hidden), b2, W3 (groups x hidden), b3.  din, hidden and
 * groups must be multiples of 16; targets must lie in [0, groups) (a row
 * whose target does not gets a NaN loss and no one-hot term).  Exposed so
 * its gradients can be checked independently. */
size_t kmMlpParamCount(int din, int hidden, int groups);
size_t kmMlpWorkspaceSize(int din, int hidden, int groups, int batch);
/* One forward + backward pass: X (batch x din BF16 bits), targets (batch),
 * P (params), G (gradients out), loss (device float, batch mean). */
kmeansCudaErrorStatus kmMlpForwardBackward(void *stream, int din, int hidden, int groups,
                                           int batch, const uint16_t *X, const int *targets,
                                           const float *P, float *G, float *loss,
                                           void *workspace, size_t workspace_bytes);

/* =========================================================================
 * 3. Legacy kernel names (KmeansCudaKernels.h:238-474), default stream.
 * Scratch arguments of the reference (Bpadded, C, Cindices) are accepted and
 * ignored: the fused kernels need none.  ClusterCenters reuses
 * centers_large as a k x (n+1) FP64 accumulator and counts_large as k int64
 * counters (both caller-zeroed per iteration, as the reference requires);
 * Compactness keeps an FP64 running total in the 8 bytes after
 * compactness[0].  The ...D forms compute in double throughout (the
 * reference's fused double argmax is an empty stub, reduction.hcu:8-11).
 * ========================================================================= */
/* ClosestCentersF/D: KmeansCudaKernels.h:238-273 */
kmeansCudaErrorStatus ClosestCentersF(const int m0, const int nRowsA, const int nColsA,
                                      const float *A, const int nColsB, const float *B,
                                      float *Bpadded, const float *normRowsOfA_squared,
                                      const float *normColsOfB_squared, const int nColsC,
                                      float *C, int *Cindices, int *CindicesFinal);
kmeansCudaErrorStatus ClosestCentersD(const int m0, const int nRowsA, const int nColsA,
                                      const double *A, const int nColsB, const double *B,
                                      double *Bpadded, const double *normRowsOfA_squared,
                                      const double *normColsOfB_squared, const int nColsC,
                                      double *C, int *Cindices, int *CindicesFinal);
/* rowTransformMinimumF/D: KmeansCudaKernels.h:290-313.  input = m x n
 * product (n = number of centres); output[m0 + i] = lowest j minimising
 * normRowsOfA_squared[i] + normColsOfB_squared[j] - 2 input[i][j] (rounded
 * as RowTransformMinimum.cu:28 does), -1 if no value is below FLT_MAX.
 * normRowsOfA_squared is indexed by the local row i, as in the reference. */
kmeansCudaErrorStatus rowTransformMinimumF(const int m0, const int m, const int n,
                                           const float *normRowsOfA_squared,
                                           const float *normColsOfB_squared, const float *input,
                                           int *output);
kmeansCudaErrorStatus rowTransformMinimumD(const int m0, const int m, const int n,
                                           const double *normRowsOfA_squared,
                                           const double *normColsOfB_squared, const double *input,
                                           int *output);
/* ClusterCentersF/D: KmeansCudaKernels.h:337-369 */
kmeansCudaErrorStatus ClusterCentersF(const int m0, const int m, const int n, const int k,
                                      const bool lastTile, const float *data, const int *indices,
                                      float *centers_large, int *counts_large, float *centers,
                                      int *counts);
kmeansCudaErrorStatus ClusterCentersD(const int m0, const int m, const int n, const int k,
                                      const bool lastTile, const double *data, const int *indices,
                                      double *centers_large, int *counts_large, double *centers,
                                      int *counts);
/* CompactnessF/D: KmeansCudaKernels.h:389-414 */
kmeansCudaErrorStatus CompactnessF(const int m0, const int m, const bool lastTile, const int n,
                                   const int k, const float *data, const int *indices,
                                   const float *centers, float *compactness);
kmeansCudaErrorStatus CompactnessD(const int m0, const int m, const bool lastTile, const int n,
                                   const int k, const double *data, const int *indices,
                                   const double *centers, double *compactness);
/* rowNormalizeF/D: KmeansCudaKernels.h:428-445; colNormalizeF/D: :458-474
 * (in = m x n = D x K) */
kmeansCudaErrorStatus rowNormalizeF(const int m0, const int m, const int n, const float *in,
                                    float *out);
kmeansCudaErrorStatus rowNormalizeD(const int m0, const int m, const int n, const double *in,
                                    double *out);
kmeansCudaErrorStatus colNormalizeF(const int m, const int n, const float *in, float *out);
kmeansCudaErrorStatus colNormalizeD(const int m, const int n, const double *in, double *out);

#ifdef __cplusplus
}
#endif

#endif /* KMEANS_AMD_H_ */
