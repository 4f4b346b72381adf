// Host engine: the Lloyd loop over the gfx950 kernels.
//
// Mirrors the reference's host class kmeans<TYPE> (main/cpp/kmeans.hpp:
// 174-658): the same public method names (buildData, initialize, compute,
// computeStandard, computeMiniBatch, closestCenters, clusterCenters,
// compactness, transpose, reset, computeTiling, copyTileToDevice,
// constructMiniBatch, getClusterCenters, getConvergenceFlag), but every step
// is stream-ordered with no per-phase host synchronisation, no __constant__
// state, and labels / compactness are returned.
#pragma once
#include <vector>

#include "km_common.hpp"

namespace km {

struct Engine {
  // problem
  long m = 0;
  int n = 0, k = 0;
  kmeansOptions opt{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int ldc = 0;

  // device buffers
  const float *dX = nullptr;
  float *dX_owned = nullptr;
  float *dC = nullptr;      // k x n
  float *dCt = nullptr;     // n x ldc
  float *dnc = nullptr;     // k
  int *dlabels = nullptr;   // m
  long long *dcounts = nullptr;
  int *dempty = nullptr;    // empty clusters of a single update call
  int *dempty_hist = nullptr;  // empty clusters, one per iteration slot
  double *dcomp = nullptr;  // one per iteration slot
  double *dq = nullptr;     // m: FP64 squared norms of the points (learned proposer only)
  float *dxn = nullptr;     // m: FP32 row norms for the library-GEMM path (resident data)
  char *dws = nullptr;
  UpdateWs ws{};
  // overlapped update when the points span more than one canonical block:
  // two side streams (default priority), their events and a second
  // per-block workspace
  hipStream_t upd_side[2] = {nullptr, nullptr};
  hipEvent_t upd_ev[7] = {};
  char *dws2 = nullptr;
  UpdateWs ws2{};
  // Resident points: the state of the last two full updates, so that an
  // update can reuse the chunk partials of the members that did not move:
  // dperm[p] (m ints: each canonical block's stable sort by label at its
  // rows), dsetup[p] (per block [cstart k | ccount k | cq k+1]), dpart[p]
  // (per block reuse_qmax x (n+1) chunk partials), dfd (per block k first
  // differing positions).  Parity reuse_cur is the latest; reuse_ok when it
  // describes a full update of the current points.
  int *dperm[2] = {nullptr, nullptr};
  int *dsetup[2] = {nullptr, nullptr};
  double *dpart[2] = {nullptr, nullptr};
  int *dfd = nullptr;
  long reuse_qmax = 0;
  int reuse_cur = 0;
  bool reuse_ok = false;
  // the last full update's sort (one of dperm; order_ok once written), the
  // order in which the next BF16/FP16 assign takes the points
  int *dorder = nullptr;
  bool order_ok = false;
  // precision path buffers (allocated on first use)
  void *mfma = nullptr;
  // candidate proposer state (options.proposer != 0)
  void *learned = nullptr;
  // library-GEMM assignment state (options.full_gemm)
  void *gemm = nullptr;

  // loop state (kmeans.cpp:1194,1235-1237)
  double criterion = -1.0;
  int maxIters = 0;
  double compactnessScore = 0.0, relErr = 1.0;
  int iters = 0;
  bool centers_set = false;
  int comp_slots = 0;
  int cur_slot = 0;               // iteration slot of the running kmeansRun
  // device-side convergence test (criterion >= 0, resident full passes):
  // ctl[0] = stop flag, ctl[1] = rows the assign kernels take (0 once
  // stopped), ctl[2] = the next slot of a graph iteration; hflag[slot] =
  // host copy of the stop flag after that slot
  int *dctl = nullptr;
  int *hflag = nullptr;
  hipEvent_t ev_iter[3] = {nullptr, nullptr, nullptr};
  const int *assign_stop = nullptr;  // dctl while such a run is enqueued
  long long last_reresolved = 0;
  kmeansStats stats{};

  // FP32 iterations over resident points in one canonical block with timing
  // off run as executable hipGraphs: the assign, the update and the
  // compactness reduction of one iteration and the commit of its compactness
  // and empty-cluster count into the slot that ctl[2] names, one graph per
  // (reuse state, parity, device-side stop).  Dropped when the points or the
  // slot arrays change; graph_ok false after a failed capture (the loop then
  // launches the kernels directly).
  hipGraphExec_t gexec[8] = {};
  bool graph_ok = true;
  double *dstage = nullptr;  // [compactness, empty count (as int), run's prev0, criterion]
  void graphs_drop();
  bool graph_eligible(bool minibatch) const;
  kmeansErrorStatus graph_iteration(bool devconv, bool *ran);

  // streamed mode: points stay in host memory and pass through HBM in tiles
  // of `tile` rows (a multiple of KM_BLOCK) on a copy stream, double-buffered
  bool streaming = false;
  long tile = 0;
  const float *hX = nullptr;
  bool host_pinned = false;
  float x_absmax = 0.0f;          // max |x| over the streamed data (FP16 scale)
  float *dXt[2] = {nullptr, nullptr};
  float *pin[2] = {nullptr, nullptr};  // staging when the host buffer is pageable
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_loaded[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};

  // mini-batch (kmeans.cpp:1253-1312)
  long mb_rows = 0;   // 0: standard loop
  long mb_index = 0;

  hipEvent_t ev[8]{};
  std::vector<hipEvent_t> iter_events;  // 5 per iteration when opt.timing

  ~Engine();
  kmeansErrorStatus create(long m_, int n_, int k_, const kmeansOptions *o);
  kmeansErrorStatus buildData(const float *data, int on_device);
  kmeansErrorStatus setCenters(const float *centers_kxn_host);
  // k sampled rows (Floyd), or with reference the rows the reference picks
  kmeansErrorStatus initialize(uint64_t seed, bool reference = false);
  kmeansErrorStatus initialize_pp(uint64_t seed);  // k-means++ seeding (resident data)
  kmeansErrorStatus reset();
  kmeansErrorStatus closestCenters(long r0, long rows);
  // empty (nullable): device counter of the empty clusters (default dempty)
  kmeansErrorStatus clusterCenters(long r0, long rows, int *empty = nullptr);
  kmeansErrorStatus compactness(long r0, long rows, int slot);
  kmeansErrorStatus transpose();
  kmeansErrorStatus constructMiniBatch(long index, long *r0, long *rows);
  kmeansErrorStatus compute();
  kmeansErrorStatus computeStandard();
  kmeansErrorStatus computeMiniBatch();
  kmeansErrorStatus getClusterCenters(float *out_kxn);
  kmeansErrorStatus getLabels(int *out);
  kmeansErrorStatus getConvergenceFlag() const {
    return relErr <= criterion ? KMEANS_SUCCESS : KMEANS_NONCONVERGENCE;
  }
  kmeansErrorStatus ensure_comp_slots(int n);
  kmeansErrorStatus run_iterations(long r0, long rows, bool minibatch);
  kmeansErrorStatus setup_streaming(const float *host);
  kmeansErrorStatus load_tile(long t, int b);
  kmeansErrorStatus stream_iteration(int slot);
  // adds the empty-cluster counts of slots [0, nslots) to the stats
  kmeansErrorStatus collect_empty(int nslots);
  void record(int slot, int phase);
  void collect_timing(int nslots);
};

}  // namespace km
