// Host engine and the extern "C" boundary (include/kmeans_amd.h).
//
// Engine::computeStandard follows main/cpp/kmeans.cpp:1186-1247 (loop while
// iters < maxIters && relErr > criterion; relErr = |comp - prev| / |comp|),
// Engine::computeMiniBatch follows kmeans.cpp:1253-1312 (each mini-batch
// replaces the centres with the means of its members, early return when
// relErr <= criterion), computeKmeansF follows kmeans.cpp:1316-1361.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <random>
#include <vector>

#include "km_engine.hpp"
#include "km_gemm.hpp"
#include "km_learned.hpp"
#include "km_mfma.hpp"
#include "km_mlp.hpp"

namespace km {

Engine::~Engine() {
  if (stream) hipStreamSynchronize(stream);
  graphs_drop();
  hipFree(dstage);
  mfma_free(this);
  learned_free(this);
  gemm_free(this);
  hipFree(dX_owned);
  hipFree(dC);
  hipFree(dCt);
  hipFree(dnc);
  hipFree(dlabels);
  hipFree(dcounts);
  hipFree(dempty);
  hipFree(dempty_hist);
  hipFree(dctl);
  hipHostFree(hflag);
  for (auto &e : ev_iter)
    if (e) hipEventDestroy(e);
  hipFree(dcomp);
  hipFree(dq);
  hipFree(dxn);
  hipFree(dws);
  hipFree(dws2);
  for (int p = 0; p < 2; ++p) {
    hipFree(dperm[p]);
    hipFree(dsetup[p]);
    hipFree(dpart[p]);
  }
  hipFree(dfd);
  for (auto &s : upd_side)
    if (s) hipStreamDestroy(s);
  for (auto &e : upd_ev)
    if (e) hipEventDestroy(e);
  for (auto &e : ev)
    if (e) hipEventDestroy(e);
  for (auto &e : iter_events)
    if (e) hipEventDestroy(e);
  for (int b = 0; b < 2; ++b) {
    hipFree(dXt[b]);
    hipHostFree(pin[b]);
    if (ev_loaded[b]) hipEventDestroy(ev_loaded[b]);
    if (ev_used[b]) hipEventDestroy(ev_used[b]);
  }
  if (copy_stream) hipStreamDestroy(copy_stream);
  if (own_stream && stream) hipStreamDestroy(stream);
}

kmeansErrorStatus Engine::create(long m_, int n_, int k_, const kmeansOptions *o) {
  if (m_ <= 0 || n_ <= 0 || k_ <= 0 || k_ > m_ || m_ > 0x7fffffffL || k_ > 16384)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  m = m_;
  n = n_;
  k = k_;
  if (o) opt = *o;
  if (opt.precision < KM_PREC_FP32 || opt.precision > KM_PREC_FP16)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (opt.full_gemm && (opt.precision != KM_PREC_FP32 || opt.proposer != KM_PROPOSER_NONE))
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (opt.precision != KM_PREC_FP32 && n_ > 256) return KMEANS_ERROR_INVALID_ARGUMENT;
  if (learned_configure(this) != KMEANS_SUCCESS) return KMEANS_ERROR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    fprintf(stderr, "kmeans_amd: no HIP device available\n");
    return KMEANS_ERROR_BUILDDATA;
  }
  KM_HIP_CHECK(hipSetDevice(opt.device), KMEANS_ERROR_BUILDDATA);
  if (opt.stream) {
    stream = (hipStream_t)opt.stream;
  } else {
    KM_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), KMEANS_ERROR_BUILDDATA);
    own_stream = true;
  }
  ldc = centers_ld(k);
  KM_HIP_CHECK(hipMalloc(&dC, sizeof(float) * (size_t)k * n), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dCt, sizeof(float) * (size_t)ldc * n), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMemsetAsync(dCt, 0, sizeof(float) * (size_t)ldc * n, stream),
               KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dnc, sizeof(float) * k), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dlabels, sizeof(int) * (size_t)m), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMemsetAsync(dlabels, 0, sizeof(int) * (size_t)m, stream), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dcounts, sizeof(long long) * k), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dempty, sizeof(int) * 2), KMEANS_ERROR_BUILDDATA);
  const size_t wsb = update_ws_layout(m, n, k, nullptr, nullptr);
  KM_HIP_CHECK(hipMalloc(&dws, wsb), KMEANS_ERROR_BUILDDATA);
  update_ws_layout(m, n, k, dws, &ws);
  if (m > KM_BLOCK && opt.tile_rows == 0) {  // resident points in several canonical blocks
    KM_HIP_CHECK(hipMalloc(&dws2, wsb), KMEANS_ERROR_BUILDDATA);
    update_ws_layout(m, n, k, dws2, &ws2);
    for (auto &s : upd_side)
      KM_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), KMEANS_ERROR_BUILDDATA);
    for (auto &e : upd_ev)
      KM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming), KMEANS_ERROR_BUILDDATA);
  }
  if (opt.tile_rows == 0) {  // resident points: the updates' reuse state
    const long nblk = km_cdiv(m, KM_BLOCK);
    reuse_qmax = std::min<long>(m, KM_BLOCK) / KM_CHUNK + k + 1;
    for (int p = 0; p < 2; ++p) {
      KM_HIP_CHECK(hipMalloc(&dperm[p], sizeof(int) * (size_t)m), KMEANS_ERROR_BUILDDATA);
      KM_HIP_CHECK(hipMalloc(&dsetup[p], sizeof(int) * (size_t)nblk * (3 * k + 1)),
                   KMEANS_ERROR_BUILDDATA);
      KM_HIP_CHECK(hipMalloc(&dpart[p], sizeof(double) * (size_t)nblk * reuse_qmax * (n + 1)),
                   KMEANS_ERROR_BUILDDATA);
    }
    KM_HIP_CHECK(hipMalloc(&dfd, sizeof(int) * (size_t)nblk * k), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dstage, sizeof(double) * 4), KMEANS_ERROR_BUILDDATA);
  }
  return ensure_comp_slots(64);
}

// Per-iteration-slot arrays of a run (compactness, empty-cluster counts and
// the host copies of the stop flag); their contents do not outlive a run.
kmeansErrorStatus Engine::ensure_comp_slots(int nslots) {
  if (nslots <= comp_slots) return KMEANS_SUCCESS;
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_BUILDDATA);
  graphs_drop();  // the graphs hold the slot arrays' addresses
  hipFree(dcomp);
  hipFree(dempty_hist);
  hipHostFree(hflag);
  dcomp = nullptr;
  dempty_hist = nullptr;
  hflag = nullptr;
  comp_slots = 0;
  KM_HIP_CHECK(hipMalloc(&dcomp, sizeof(double) * nslots), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&dempty_hist, sizeof(int) * nslots), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipHostMalloc(&hflag, sizeof(int) * nslots, hipHostMallocDefault),
               KMEANS_ERROR_BUILDDATA);
  comp_slots = nslots;
  if (!dctl) {
    KM_HIP_CHECK(hipMalloc(&dctl, sizeof(int) * 4), KMEANS_ERROR_BUILDDATA);
    for (auto &e : ev_iter)
      KM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming), KMEANS_ERROR_BUILDDATA);
  }
  return KMEANS_SUCCESS;
}

// kmeans.cpp:540-695 (buildData) without the row shuffle of :554-564.
kmeansErrorStatus Engine::buildData(const float *data, int on_device) {
  if (!data) return KMEANS_ERROR_INVALID_ARGUMENT;
  // the BF16/FP16 assign's side-stream work may still read the old points
  if (mfma) KM_HIP_CHECK(mfma_pair_join(this), KMEANS_ERROR_BUILDDATA);
  reuse_ok = false;  // new points: no chunk partials to reuse
  graphs_drop();     // the graphs hold the old points' address
  if (!on_device && opt.tile_rows > 0 && opt.tile_rows < m) return setup_streaming(data);
  if (on_device) {
    dX = data;
  } else {
    if (!dX_owned)
      KM_HIP_CHECK(hipMalloc(&dX_owned, sizeof(float) * (size_t)m * n), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMemcpyAsync(dX_owned, data, sizeof(float) * (size_t)m * n,
                                hipMemcpyHostToDevice, stream),
                 KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_BUILDDATA);
    dX = dX_owned;
  }
  // per-point FP64 squared norms for the learned proposer's certificates
  // (the update takes its squared-norm moments from the rows it reads)
  if (learned) {
    if (!dq) KM_HIP_CHECK(hipMalloc(&dq, sizeof(double) * (size_t)m), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(launch_point_sq(stream, m, n, dX, dq), KMEANS_ERROR_BUILDDATA);
  }
  // the GEMM path's row norms, once per data set (k_row_norms, as per tile)
  if (opt.full_gemm) {
    if (!dxn) KM_HIP_CHECK(hipMalloc(&dxn, sizeof(float) * (size_t)m), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(launch_row_norms(stream, m, n, dX, dxn), KMEANS_ERROR_BUILDDATA);
  }
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_BUILDDATA);
  if (mfma) mfma_invalidate(this);
  if (opt.precision != KM_PREC_FP32) KM_HIP_CHECK(mfma_prepare(this), KMEANS_ERROR_BUILDDATA);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::setCenters(const float *centers) {
  if (!centers) return KMEANS_ERROR_INVALID_ARGUMENT;
  KM_HIP_CHECK(hipMemcpyAsync(dC, centers, sizeof(float) * (size_t)k * n, hipMemcpyHostToDevice,
                              stream),
               KMEANS_ERROR_INITIALIZE);
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_INITIALIZE);
  centers_set = true;
  relErr = 1.0;
  compactnessScore = 0.0;
  return transpose();
}

// Centre norms and the D x ldc copy the assign kernels read (replaces the
// per-iteration cuBLAS geam transpose of kmeans.cpp:1109-1139).
kmeansErrorStatus Engine::transpose() {
  KM_HIP_CHECK(launch_transpose(stream, k, n, dC, dCt, ldc), KMEANS_ERROR_TRANSPOSE);
  KM_HIP_CHECK(launch_col_norms(stream, k, n, dC, dnc), KMEANS_ERROR_TRANSPOSE);
  mfma_centers_changed(this, true);
  if (opt.precision != KM_PREC_FP32) KM_HIP_CHECK(mfma_prepare(this), KMEANS_ERROR_TRANSPOSE);
  return KMEANS_SUCCESS;
}

// The reference's initial centres without reordering the data.  Its
// buildData stores row indices[i] of the caller's data as row i, with
// indices = 0..m-1 shuffled by std::shuffle with std::default_random_engine(0)
// (kmeans.cpp:553-556); initialize then takes rows points[0..k) of that
// shuffled copy, points[i] = rand() % m after srand(42), redrawn while
// std::binary_search over the points so far (an unsorted vector, so some
// duplicates get through) finds it (kmeans.cpp:876-883).  Centre i is
// therefore caller row indices[points[i]].  The same library routines are
// called here; glibc's rand() is reproduced with a private random_r state of
// the same type (128 bytes, seeded 42), so the caller's rand() sequence is
// not disturbed.
void reference_init_rows(long m, int k, int *rows) {
  std::vector<int> indices(m);
  for (long i = 0; i < m; ++i) indices[i] = (int)i;
  std::shuffle(indices.begin(), indices.end(), std::default_random_engine(0));
  char state[128];
  struct random_data rd;
  memset(&rd, 0, sizeof(rd));
  memset(state, 0, sizeof(state));
  initstate_r(42, state, sizeof(state), &rd);
  auto next = [&]() {
    int32_t r = 0;
    random_r(&rd, &r);
    return (int)(r % (int)m);
  };
  std::vector<int> points;
  points.reserve(k);
  for (int i = 0; i < k; ++i) {
    int index = next();
    while (std::binary_search(points.begin(), points.end(), index)) index = next();
    points.push_back(index);
  }
  for (int i = 0; i < k; ++i) rows[i] = indices[points[i]];
}

// kmeans.cpp:822-907: k rows as initial centres.  reference: the rows the
// reference picks (reference_init_rows); otherwise k distinct indices drawn
// with Floyd's algorithm from the counter-based generator (seed, stream 4).
kmeansErrorStatus Engine::initialize(uint64_t seed, bool reference) {
  if (!dX && !hX) return KMEANS_ERROR_INITIALIZE;
  std::vector<int> idx(k);
  if (reference)
    reference_init_rows(m, k, idx.data());
  else
    sample_rows(seed, m, k, idx.data());
  if (streaming) {  // gather the rows on the host
    std::vector<float> c((size_t)k * n);
    for (int j = 0; j < k; ++j)
      memcpy(&c[(size_t)j * n], hX + (size_t)idx[j] * n, sizeof(float) * n);
    return setCenters(c.data());
  }
  int *didx = nullptr;
  KM_HIP_CHECK(hipMalloc(&didx, sizeof(int) * k), KMEANS_ERROR_INITIALIZE);
  hipError_t e = hipMemcpyAsync(didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = launch_gather_rows(stream, k, n, dX, didx, dC);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  hipFree(didx);
  KM_HIP_CHECK(e, KMEANS_ERROR_INITIALIZE);
  centers_set = true;
  relErr = 1.0;
  compactnessScore = 0.0;
  return transpose();
}

// k-means++ seeding of the resident points (no counterpart in the
// reference, whose initAlgorithm is ignored): the picks on the GPU, then the
// picked rows as the centres.
kmeansErrorStatus Engine::initialize_pp(uint64_t seed) {
  if (!dX) return KMEANS_ERROR_INITIALIZE;
  int *didx = nullptr;
  KM_HIP_CHECK(hipMalloc(&didx, sizeof(int) * k), KMEANS_ERROR_INITIALIZE);
  hipError_t e = kmeanspp(stream, m, n, k, dX, seed, didx);
  if (e == hipSuccess) e = launch_gather_rows(stream, k, n, dX, didx, dC);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  hipFree(didx);
  KM_HIP_CHECK(e, KMEANS_ERROR_INITIALIZE);
  centers_set = true;
  relErr = 1.0;
  compactnessScore = 0.0;
  return transpose();
}

// kmeans.cpp:1143-1170: nothing to clear (no G x K x D slabs); the update
// zeroes its own FP64 accumulators.
kmeansErrorStatus Engine::reset() { return KMEANS_SUCCESS; }

kmeansErrorStatus Engine::closestCenters(long r0, long rows) {
  hipError_t e;
  if (opt.full_gemm) {
    e = gemm_assign(this, dX + r0 * n, dxn ? dxn + r0 : nullptr, rows, dlabels + r0);
  } else if (opt.precision == KM_PREC_FP32) {
    // under a device-side convergence test the row count comes from ctl[1]
    // (0 after the converged iteration)
    e = launch_assign_f32(stream, rows, assign_stop ? assign_stop + 1 : nullptr, n, dX + r0 * n,
                          nullptr, k, dCt, ldc, dnc, dlabels + r0, nullptr);
  } else {
    e = mfma_assign(this, r0, rows);
  }
  KM_HIP_CHECK(e, KMEANS_ERROR_CLOSEST_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::clusterCenters(long r0, long rows, int *empty) {
  int *dempty = empty ? empty : this->dempty;
  // every label final: the BF16/FP16 assign's two-candidate decisions may
  // still run on its side stream
  if (mfma) KM_HIP_CHECK(mfma_pair_join(this), KMEANS_ERROR_CLUSTER_CENTERS);
  ws.perm_out = nullptr;
  ws2.perm_out = nullptr;
  // a full update over all resident points sorts each canonical block into
  // the other parity of the reuse state, reusing the chunk partials of the
  // last full update wherever a centre's members did not change; its sort is
  // the order of the next assign
  const bool full = dperm[0] && r0 == 0 && rows == m;
  std::vector<UpdateWs> views;
  const int nw = reuse_cur ^ 1;
  if (full) {
    const long nblk = km_cdiv(m, KM_BLOCK);
    const size_t ps = (size_t)reuse_qmax * (n + 1);
    views.resize(nblk);
    for (long b = 0; b < nblk; ++b) {
      UpdateWs v = ((b & 1) && upd_side[0]) ? ws2 : ws;
      v.perm_out = dperm[nw];
      v.cstart = dsetup[nw] + (size_t)b * (3 * k + 1);
      v.ccount = v.cstart + k;
      v.cq = v.ccount + k;
      v.P = dpart[nw] + b * ps;
      v.fd = dfd + (size_t)b * k;
      if (reuse_ok) {
        v.perm_prev = dperm[reuse_cur];
        v.cs_prev = dsetup[reuse_cur] + (size_t)b * (3 * k + 1);
        v.P_prev = dpart[reuse_cur] + b * ps;
      }
      views[b] = v;
    }
  }
  if (rows > KM_BLOCK && upd_side[0]) {
    // several canonical blocks: sorts and chunk sums of alternate blocks on
    // two side streams, the in-order additions on the engine's stream
    const UpdateOverlap ov{{upd_side[0], upd_side[1]}, upd_ev[0], {upd_ev[1], upd_ev[2]},
                           {upd_ev[3], upd_ev[4]}, {upd_ev[5], upd_ev[6]}, {&ws, &ws2}};
    KM_HIP_CHECK(launch_update_zero(stream, n, k, ws.S, ws.cnt, dempty),
                 KMEANS_ERROR_CLUSTER_CENTERS);
    KM_HIP_CHECK(launch_update_accumulate_2s(stream, ov, rows, n, k, dX + r0 * n, dlabels + r0,
                                             ws.S, ws.cnt, full ? views.data() : nullptr),
                 KMEANS_ERROR_CLUSTER_CENTERS);
    KM_HIP_CHECK(launch_update_finalize(stream, n, k, ws.S, ws.cnt, dC, dC, dCt, ldc, dnc, dcounts,
                                        dempty, ws.compj),
                 KMEANS_ERROR_CLUSTER_CENTERS);
  } else {
    KM_HIP_CHECK(launch_update(stream, rows, n, k, dX + r0 * n, dlabels + r0, dC, dC, dCt, ldc,
                               dnc, dcounts, dempty, full ? views[0] : ws),
                 KMEANS_ERROR_CLUSTER_CENTERS);
  }
  if (full) {
    reuse_cur = nw;
    reuse_ok = true;
    dorder = dperm[nw];
    order_ok = true;
  } else {
    reuse_ok = false;
  }
  if (opt.precision != KM_PREC_FP32) KM_HIP_CHECK(mfma_centers(this), KMEANS_ERROR_CLUSTER_CENTERS);
  return KMEANS_SUCCESS;
}

// kmeans.cpp:1045-1104: compactness against the new centres.  The update
// already produced the per-centre terms from its FP64 moments (no second
// pass over the points); this reduces them in the canonical order.
kmeansErrorStatus Engine::compactness(long r0, long rows, int slot) {
  (void)r0;
  (void)rows;
  KM_HIP_CHECK(launch_comp_reduce(stream, k, ws.compj, dcomp + slot), KMEANS_ERROR_COMPACTNESS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::constructMiniBatch(long index, long *r0, long *rows) {
  // windows [0,mb), [mb,2mb), ..., [q*mb, m), then back to row 0
  const long per_epoch = (m + mb_rows - 1) / mb_rows;
  const long start = (index % per_epoch) * mb_rows;
  *r0 = start;
  *rows = std::min(mb_rows, m - start);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::compute() {
  if ((!dX && !streaming) || !centers_set) return KMEANS_ERROR_COMPUTE;
  return mb_rows > 0 ? computeMiniBatch() : computeStandard();
}

void Engine::record(int slot, int phase) {
  if (!opt.timing) return;
  const size_t idx = (size_t)slot * 5 + phase;
  while (iter_events.size() <= idx) {
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    iter_events.push_back(e);
  }
  hipEventRecord(iter_events[idx], stream);
}

void Engine::collect_timing(int nslots) {
  if (!opt.timing) return;
  hipStreamSynchronize(stream);
  for (int s = 0; s < nslots; ++s) {
    float t[4] = {0, 0, 0, 0};
    for (int p = 0; p < 4; ++p)
      hipEventElapsedTime(&t[p], iter_events[s * 5 + p], iter_events[s * 5 + p + 1]);
    stats.ms_assign += t[0];
    stats.ms_reresolve += t[1];
    stats.ms_update += t[2];
    stats.ms_compactness += t[3];
    stats.ms_total += t[0] + t[1] + t[2] + t[3];
  }
}

// Control words of a run: no stop, all rows assigned (ctl[1], read by the
// assign kernels under a device-side convergence test), slot 0 next
// (ctl[2], read by the graph iterations' commit); with stage, the run's
// convergence parameters for that commit (stage[2] = the compactness before
// the run, stage[3] = the criterion).
__global__ void k_ctl_reset(int *ctl, int rows, double *stage, double prev0, double criterion) {
  ctl[0] = 0;
  ctl[1] = rows;
  ctl[2] = 0;
  if (stage) {
    stage[2] = prev0;
    stage[3] = criterion;
  }
}

// The reference's convergence test (kmeans.cpp:1194,1235-1237) after
// iteration slot s: relErr = |comp - prev| / |comp| with prev the previous
// slot's compactness (prev0 before slot 0); the loop continues only while
// relErr > criterion.  Once it stops, ctl[1] = 0 gives the assign kernels of
// the iterations already enqueued behind it no rows, so the labels stay
// those of the converged iteration and the rest of such an iteration
// recomputes the same centres and compactness from them.
__global__ void k_converge(const double *comp, int s, double prev0, double criterion, int *ctl) {
  if (ctl[0]) return;
  const double c = comp[s], prev = s > 0 ? comp[s - 1] : prev0;
  const double rel = fabs(c - prev) / fabs(c);
  if (!(rel > criterion)) {
    ctl[0] = 1;
    ctl[1] = 0;
  }
}

// The last node of a graph iteration: its compactness and empty-cluster
// count (stage[0], and stage[1] read as an int) into slot s = ctl[2], the
// next slot in ctl[2]; with test, the convergence test of k_converge for
// that slot against stage[2] (before slot 0) and criterion stage[3].
__global__ void k_slot_commit(const double *stage, double *comp, int *empty_hist, int *ctl,
                              int test) {
  const int s = ctl[2];
  ctl[2] = s + 1;
  comp[s] = stage[0];
  empty_hist[s] = *reinterpret_cast<const int *>(stage + 1);
  if (!test || ctl[0]) return;
  const double c = comp[s], prev = s > 0 ? comp[s - 1] : stage[2];
  const double rel = fabs(c - prev) / fabs(c);
  if (!(rel > stage[3])) {
    ctl[0] = 1;
    ctl[1] = 0;
  }
}

void Engine::graphs_drop() {
  bool any = false;
  for (auto &g : gexec) any |= g != nullptr;
  if (!any) return;
  hipStreamSynchronize(stream);  // no launch of them still in flight
  for (auto &g : gexec) {
    if (g) hipGraphExecDestroy(g);
    g = nullptr;
  }
}

bool Engine::graph_eligible(bool minibatch) const {
  return graph_ok && !opt.timing && opt.precision == KM_PREC_FP32 && !opt.full_gemm && !learned &&
         !streaming && !minibatch && dperm[0] && dstage && m <= KM_BLOCK;
}

// One full-pass iteration (assign, update, compactness into dstage and the
// slot commit) as a graph launch; the graph of this state is captured from
// the engine's own launch sequence on first use.  *ran false: no graph
// (capture or instantiation failed; graph_ok is cleared) and nothing was
// enqueued.
kmeansErrorStatus Engine::graph_iteration(bool devconv, bool *ran) {
  *ran = false;
  // after a failed capture no graph runs again: the commit's device-side
  // slot counter only advances in graph iterations
  if (!graph_ok) return KMEANS_SUCCESS;
  const int nw = reuse_cur ^ 1;
  const int v = (reuse_ok ? 1 : 0) | (nw << 1) | (devconv ? 4 : 0);
  if (!gexec[v]) {
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      graph_ok = false;
      return KMEANS_SUCCESS;
    }
    // clusterCenters advances the reuse state as it enqueues; the launch
    // below advances it again, so it is put back after the capture
    const int rc = reuse_cur;
    const bool ro = reuse_ok, ook = order_ok;
    int *dord = dorder;
    kmeansErrorStatus err = closestCenters(0, m);
    if (err == KMEANS_SUCCESS)
      err = clusterCenters(0, m, reinterpret_cast<int *>(dstage + 1));
    if (err == KMEANS_SUCCESS && launch_comp_reduce(stream, k, ws.compj, dstage) != hipSuccess)
      err = KMEANS_ERROR_COMPACTNESS;
    if (err == KMEANS_SUCCESS) {
      hipLaunchKernelGGL(k_slot_commit, dim3(1), dim3(1), 0, stream, dstage, dcomp, dempty_hist,
                         dctl, devconv ? 1 : 0);
      if (hipGetLastError() != hipSuccess) err = KMEANS_ERROR_COMPUTE;
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(stream, &g);
    reuse_cur = rc;
    reuse_ok = ro;
    order_ok = ook;
    dorder = dord;
    hipGraphExec_t x = nullptr;
    const bool ok = err == KMEANS_SUCCESS && ce == hipSuccess && g &&
                    hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) {
      (void)hipGetLastError();
      if (x) hipGraphExecDestroy(x);
      graph_ok = false;
      return KMEANS_SUCCESS;
    }
    gexec[v] = x;
  }
  KM_HIP_CHECK(hipGraphLaunch(gexec[v], stream), KMEANS_ERROR_COMPUTE);
  // the host side of a full update (clusterCenters)
  reuse_cur = nw;
  reuse_ok = true;
  dorder = dperm[nw];
  order_ok = true;
  ++stats.graph_launches;
  *ran = true;
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::collect_empty(int nslots) {
  if (nslots <= 0) return KMEANS_SUCCESS;
  std::vector<int> emp(nslots);
  KM_HIP_CHECK(hipMemcpyAsync(emp.data(), dempty_hist, sizeof(int) * nslots, hipMemcpyDeviceToHost,
                              stream),
               KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_COMPUTE);
  for (int s = 0; s < nslots; ++s) stats.empty_clusters += emp[s];
  return KMEANS_SUCCESS;
}

// The loop of kmeans.cpp:1186-1247.  Three ways to run it:
//  - criterion < 0 (a fixed iteration count): everything enqueued, the
//    compactness history read once at the end;
//  - criterion >= 0 over resident points in full passes: the convergence
//    test runs on the device after each iteration (k_converge); the host
//    keeps at most three iterations in flight and stops enqueuing once it
//    sees the flag, then finds the converged iteration in the history;
//  - otherwise (mini-batch windows, streamed tiles, the proposer, the
//    library GEMM) one host round trip per iteration.
// FP32 full passes over resident points in one canonical block, timing off,
// run each iteration as one graph launch (graph_iteration); the labels,
// centres and compactness are those of the launches the graph was captured
// from.
// Statistics count iterations up to the converged one.
kmeansErrorStatus Engine::run_iterations(long r0, long rows, bool minibatch) {
  kmeansErrorStatus err;
  const bool crit = criterion >= 0.0;
  const bool devconv = crit && !streaming && !minibatch && !learned && !opt.full_gemm;
  const bool sync_each = crit && !devconv;
  int it0 = iters, slot = 0;
  if ((err = ensure_comp_slots(std::max(maxIters, 1))) != KMEANS_SUCCESS) return err;
  const double prev0 = compactnessScore;
  const bool graphs = graph_eligible(minibatch);
  if (devconv || graphs) {
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, stream, dctl, (int)rows,
                       graphs ? dstage : nullptr, prev0, criterion);
    KM_LAUNCH_CHECK(KMEANS_ERROR_COMPUTE);
  }
  if (devconv) assign_stop = dctl;
  // assign_stop is cleared however the loop is left
  struct StopGuard {
    const int **p;
    ~StopGuard() { *p = nullptr; }
  } guard{&assign_stop};
  // one iteration's compactness and empty-cluster count back to the host
  auto read_slot = [&](int s, double *comp, int *emp) -> kmeansErrorStatus {
    KM_HIP_CHECK(hipMemcpyAsync(comp, dcomp + s, sizeof(double), hipMemcpyDeviceToHost, stream),
                 KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipMemcpyAsync(emp, dempty_hist + s, sizeof(int), hipMemcpyDeviceToHost, stream),
                 KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_COMPUTE);
    return KMEANS_SUCCESS;
  };
  while (iters < maxIters && relErr > criterion) {
    bool graph_ran = false;
    long b0 = r0, br = rows;
    if (minibatch) constructMiniBatch(mb_index++, &b0, &br);
    cur_slot = slot;
    if (streaming) {
      record(slot, 0);
      record(slot, 1);
      record(slot, 2);
      if ((err = stream_iteration(slot)) != KMEANS_SUCCESS) return err;
      record(slot, 3);
      record(slot, 4);
    } else {
      if (learned && minibatch) return KMEANS_ERROR_INVALID_ARGUMENT;
      if (learned_active(this)) {
        // proposer iteration: the proposal and the grouped exact scores are
        // the assign phase, the full search of uncertified points the
        // re-resolve
        if ((err = learned_prepare(this)) != KMEANS_SUCCESS) return err;
        if (mfma) mfma_bounds_invalidate(this);
        record(slot, 0);
        KM_HIP_CHECK(learned_assign(this), KMEANS_ERROR_CLOSEST_CENTERS);
        record(slot, 1);
        KM_HIP_CHECK(learned_fallback(this), KMEANS_ERROR_CLOSEST_CENTERS);
        record(slot, 2);
      } else {
        if (graphs && b0 == 0 && br == m &&
            (err = graph_iteration(devconv, &graph_ran)) != KMEANS_SUCCESS)
          return err;
        if (!graph_ran) {
          record(slot, 0);
          if ((err = closestCenters(b0, br)) != KMEANS_SUCCESS) return err;
          record(slot, 1);
          if (opt.precision != KM_PREC_FP32)
            KM_HIP_CHECK(mfma_reresolve(this, b0, br), KMEANS_ERROR_CLOSEST_CENTERS);
          record(slot, 2);
          if (learned) KM_HIP_CHECK(learned_collect(this, iters), KMEANS_ERROR_CLOSEST_CENTERS);
        }
      }
      if (!graph_ran) {
        if ((err = clusterCenters(b0, br, dempty_hist + slot)) != KMEANS_SUCCESS) return err;
        record(slot, 3);
        if ((err = compactness(b0, br, slot)) != KMEANS_SUCCESS) return err;
        record(slot, 4);
      }
    }
    if (devconv) {
      if (!graph_ran) {  // (the graph iteration's commit made the test)
        hipLaunchKernelGGL(k_converge, dim3(1), dim3(1), 0, stream, dcomp, slot, prev0, criterion,
                           dctl);
        KM_LAUNCH_CHECK(KMEANS_ERROR_COMPUTE);
      }
      KM_HIP_CHECK(hipMemcpyAsync(hflag + slot, dctl, sizeof(int), hipMemcpyDeviceToHost, stream),
                   KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(hipEventRecord(ev_iter[slot % 3], stream), KMEANS_ERROR_COMPUTE);
    }
    ++iters;
    ++slot;
    if (sync_each) {
      int emp = 0;
      double comp = 0.0;
      if ((err = read_slot(slot - 1, &comp, &emp)) != KMEANS_SUCCESS) return err;
      stats.empty_clusters += emp;
      const double prev = compactnessScore;
      compactnessScore = comp;
      relErr = fabs(comp - prev) / fabs(comp);
      if (minibatch && relErr <= criterion) break;
    } else {
      // no host round trip inside the loop; with the device-side test the
      // host waits for the iteration two before the one just enqueued (two
      // stay queued, so the GPU does not idle) and stops enqueuing once that
      // one had converged
      relErr = 1.0 + fabs(criterion);
      if (devconv && slot >= 3) {
        KM_HIP_CHECK(hipEventSynchronize(ev_iter[(slot - 3) % 3]), KMEANS_ERROR_COMPUTE);
        if (hflag[slot - 3]) break;
      }
    }
  }
  int used = slot;  // iterations that count (up to the converged one)
  if (!sync_each && slot > 0) {
    std::vector<double> hist(slot);
    KM_HIP_CHECK(hipMemcpyAsync(hist.data(), dcomp, sizeof(double) * slot, hipMemcpyDeviceToHost,
                                stream),
                 KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_COMPUTE);
    for (int s = 0; s < slot; ++s) {
      const double prev = compactnessScore;
      compactnessScore = hist[s];
      relErr = fabs(hist[s] - prev) / fabs(hist[s]);
      if (devconv && !(relErr > criterion)) {  // the test k_converge made
        used = s + 1;
        break;
      }
    }
    iters = it0 + used;
    if ((err = collect_empty(used)) != KMEANS_SUCCESS) return err;
  }
  if (opt.precision != KM_PREC_FP32) mfma_collect(this, slot);  // later slots add 0
  learned_stats(this);
  collect_timing(used);
  stats.iterations = iters;
  stats.compactness = compactnessScore;
  stats.rel_err = relErr;
  return KMEANS_SUCCESS;
}

// Host-resident points streamed through HBM (kmeans.cpp:302-536 tiling and
// :700-789 copyTileToDevice, here asynchronous and double-buffered).  Tiles
// are multiples of KM_BLOCK rows, so the canonical sums equal the resident
// path's; the compactness comes from the update's moments, so each
// iteration reads the points exactly once.
kmeansErrorStatus Engine::setup_streaming(const float *host) {
  if (opt.tile_rows % KM_BLOCK != 0) {
    fprintf(stderr, "kmeans_amd: tile_rows must be a multiple of KM_BLOCK (%ld)\n", (long)KM_BLOCK);
    return KMEANS_ERROR_TILING;
  }
  streaming = true;
  tile = opt.tile_rows;
  hX = host;
  hipPointerAttribute_t attr;
  host_pinned = hipPointerGetAttributes(&attr, host) == hipSuccess &&
                attr.type == hipMemoryTypeHost;
  (void)hipGetLastError();  // a pageable pointer leaves an error code behind
  for (int b = 0; b < 2; ++b) {
    KM_HIP_CHECK(hipMalloc(&dXt[b], sizeof(float) * (size_t)tile * n), KMEANS_ERROR_TILING);
    if (!host_pinned)
      KM_HIP_CHECK(hipHostMalloc(&pin[b], sizeof(float) * (size_t)tile * n, hipHostMallocDefault),
                   KMEANS_ERROR_TILING);
    KM_HIP_CHECK(hipEventCreateWithFlags(&ev_loaded[b], hipEventDisableTiming), KMEANS_ERROR_TILING);
    KM_HIP_CHECK(hipEventCreateWithFlags(&ev_used[b], hipEventDisableTiming), KMEANS_ERROR_TILING);
  }
  KM_HIP_CHECK(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking), KMEANS_ERROR_TILING);
  // max |x| for the FP16 scale: one host pass
  if (opt.precision == KM_PREC_FP16) {
    float mx = 0.0f;
    const size_t tot = (size_t)m * n;
    for (size_t i = 0; i < tot; ++i) mx = fmaxf(mx, fabsf(host[i]));
    x_absmax = mx;
  }
  return KMEANS_SUCCESS;
}

// Issue the copy of tile t into buffer b on the copy stream.
kmeansErrorStatus Engine::load_tile(long t, int b) {
  const long r0 = t * tile, rows = std::min(tile, m - r0);
  KM_HIP_CHECK(hipStreamWaitEvent(copy_stream, ev_used[b], 0), KMEANS_ERROR_COPY_TILE);
  const float *src = hX + r0 * n;
  if (!host_pinned) {
    // the previous copy out of this staging buffer must be finished
    KM_HIP_CHECK(hipEventSynchronize(ev_loaded[b]), KMEANS_ERROR_COPY_TILE);
    memcpy(pin[b], src, sizeof(float) * (size_t)rows * n);
    src = pin[b];
  }
  KM_HIP_CHECK(hipMemcpyAsync(dXt[b], src, sizeof(float) * (size_t)rows * n, hipMemcpyHostToDevice,
                              copy_stream),
               KMEANS_ERROR_COPY_TILE);
  KM_HIP_CHECK(hipEventRecord(ev_loaded[b], copy_stream), KMEANS_ERROR_COPY_TILE);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::stream_iteration(int slot) {
  kmeansErrorStatus err;
  const long nt = (m + tile - 1) / tile;
  int *empty = dempty_hist + slot;
  KM_HIP_CHECK(launch_update_zero(stream, n, k, ws.S, ws.cnt, empty), KMEANS_ERROR_CLUSTER_CENTERS);
  // the first tile may reuse buffer 0 only after the previous iteration used it
  if ((err = load_tile(0, 0)) != KMEANS_SUCCESS) return err;
  for (long t = 0; t < nt; ++t) {
    const int b = (int)(t & 1);
    const long r0 = t * tile, rows = std::min(tile, m - r0);
    if (t + 1 < nt)
      if ((err = load_tile(t + 1, b ^ 1)) != KMEANS_SUCCESS) return err;
    KM_HIP_CHECK(hipStreamWaitEvent(stream, ev_loaded[b], 0), KMEANS_ERROR_COPY_TILE);
    if (opt.full_gemm) {
      KM_HIP_CHECK(gemm_assign(this, dXt[b], nullptr, rows, dlabels + r0),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    } else if (opt.precision == KM_PREC_FP32) {
      KM_HIP_CHECK(launch_assign_f32(stream, rows, nullptr, n, dXt[b], nullptr, k, dCt, ldc, dnc,
                                     dlabels + r0, nullptr),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    } else {
      KM_HIP_CHECK(mfma_prepare_rows(this, dXt[b], rows), KMEANS_ERROR_CLOSEST_CENTERS);
      KM_HIP_CHECK(mfma_assign_view(this, dXt[b], mfma_tile_xh(this), mfma_tile_xnorm(this),
                                    dlabels + r0, rows),
                   KMEANS_ERROR_CLOSEST_CENTERS);
      KM_HIP_CHECK(mfma_reresolve_view(this, dXt[b], mfma_tile_xh(this), mfma_tile_xnorm(this),
                                       dlabels + r0, rows),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    }
    KM_HIP_CHECK(launch_update_accumulate(stream, rows, n, k, dXt[b], dlabels + r0, ws.S, ws.cnt,
                                          ws),
                 KMEANS_ERROR_CLUSTER_CENTERS);
    KM_HIP_CHECK(hipEventRecord(ev_used[b], stream), KMEANS_ERROR_COPY_TILE);
  }
  KM_HIP_CHECK(launch_update_finalize(stream, n, k, ws.S, ws.cnt, dC, dC, dCt, ldc, dnc, dcounts,
                                      empty, ws.compj),
               KMEANS_ERROR_CLUSTER_CENTERS);
  if (opt.precision != KM_PREC_FP32) KM_HIP_CHECK(mfma_centers(this), KMEANS_ERROR_CLUSTER_CENTERS);
  KM_HIP_CHECK(launch_comp_reduce(stream, k, ws.compj, dcomp + slot), KMEANS_ERROR_COMPACTNESS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::computeStandard() { return run_iterations(0, m, false); }
kmeansErrorStatus Engine::computeMiniBatch() { return run_iterations(0, m, true); }

kmeansErrorStatus Engine::getClusterCenters(float *out) {
  KM_HIP_CHECK(hipMemcpyAsync(out, dC, sizeof(float) * (size_t)k * n, hipMemcpyDeviceToHost, stream),
               KMEANS_ERROR_GET_CENTERS);
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_GET_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus Engine::getLabels(int *out) {
  if (mfma) KM_HIP_CHECK(mfma_pair_join(this), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipMemcpyAsync(out, dlabels, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost, stream),
               KMEANS_ERROR_GET_CENTERS);
  KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_GET_CENTERS);
  return KMEANS_SUCCESS;
}

// ---------------------------------------------------------------- shared
// Workspace for the legacy kernel names (the reference's signatures carry no
// workspace argument); grown on demand, kept until process exit.
static std::mutex g_legacy_mu;
static char *g_legacy_ws = nullptr;
static size_t g_legacy_bytes = 0;
static void *g_legacy_ct = nullptr;
static size_t g_legacy_ct_bytes = 0;

static hipError_t legacy_ws(long n, int d, int k, UpdateWs *ws) {
  const size_t need = update_ws_layout(n, d, k, nullptr, nullptr);
  if (need > g_legacy_bytes) {
    hipFree(g_legacy_ws);
    g_legacy_ws = nullptr;
    g_legacy_bytes = 0;
    hipError_t e = hipMalloc(&g_legacy_ws, need);
    if (e != hipSuccess) return e;
    g_legacy_bytes = need;
  }
  update_ws_layout(n, d, k, g_legacy_ws, ws);
  return hipSuccess;
}

}  // namespace km

using km::Engine;

struct kmeansEngine {
  Engine e;
};

// ====================================================================== C API
extern "C" {

const char *kmeansGetErrorString(kmeansErrorStatus err) {
  switch (err) {
    case KMEANS_ERROR_BUILDDATA: return "error in building the data";
    case KMEANS_ERROR_INITIALIZE: return "error in the initialization";
    case KMEANS_ERROR_CLOSEST_CENTERS: return "error in computing the closest centers";
    case KMEANS_ERROR_CLUSTER_CENTERS: return "error in computing the cluster centers";
    case KMEANS_ERROR_COMPACTNESS: return "error in computing the compactness";
    case KMEANS_ERROR_TRANSPOSE: return "error in transpose";
    case KMEANS_ERROR_RESET: return "error in reset";
    case KMEANS_ERROR_COMPUTE: return "error in compute";
    case KMEANS_ERROR_GET_CENTERS: return "error in copying the data back to the host";
    case KMEANS_ERROR_TILING: return "error in the tiling";
    case KMEANS_ERROR_COPY_TILE: return "error in copy tile";
    case KMEANS_ERROR_CONSTRUCT_MINIBATCH: return "error in constructing the minibatch";
    case KMEANS_NONCONVERGENCE: return "algorithm did not converge";
    case KMEANS_ERROR_INVALID_ARGUMENT: return "invalid argument";
    default: return "no error";
  }
}

kmeansErrorStatus kmeansCreate(kmeansHandle *h, long m, int n, int k, const kmeansOptions *opt) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  *h = nullptr;
  kmeansEngine *p = new (std::nothrow) kmeansEngine();
  if (!p) return KMEANS_ERROR_BUILDDATA;
  kmeansErrorStatus st = p->e.create(m, n, k, opt);
  if (st != KMEANS_SUCCESS) {
    delete p;
    return st;
  }
  *h = p;
  return KMEANS_SUCCESS;
}

void kmeansDestroy(kmeansHandle h) { delete h; }

kmeansErrorStatus kmeansSetData(kmeansHandle h, const float *data, int on_device) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.buildData(data, on_device);
}

kmeansErrorStatus kmeansSetCenters(kmeansHandle h, const float *centers) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.setCenters(centers);
}

kmeansErrorStatus kmeansInitPlusPlus(kmeansHandle h, uint64_t seed) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.initialize_pp(seed);
}

kmeansErrorStatus kmeansRun(kmeansHandle h, int maxIters, double criterion) {
  if (!h || maxIters < 0) return KMEANS_ERROR_INVALID_ARGUMENT;
  h->e.maxIters = h->e.iters + maxIters;
  h->e.criterion = criterion;
  if (h->e.relErr <= criterion) h->e.relErr = 1.0 + fabs(criterion);
  kmeansErrorStatus st = h->e.compute();
  if (st != KMEANS_SUCCESS) return st;
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansGetCenters(kmeansHandle h, float *centers) {
  if (!h || !centers) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.getClusterCenters(centers);
}

kmeansErrorStatus kmeansGetLabels(kmeansHandle h, int *labels) {
  if (!h || !labels) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.getLabels(labels);
}

kmeansErrorStatus kmeansGetStats(kmeansHandle h, kmeansStats *stats) {
  if (!h || !stats) return KMEANS_ERROR_INVALID_ARGUMENT;
  *stats = h->e.stats;
  return KMEANS_SUCCESS;
}

const float *kmeansDeviceCenters(kmeansHandle h) { return h ? h->e.dC : nullptr; }
const int *kmeansDeviceLabels(kmeansHandle h) { return h ? h->e.dlabels : nullptr; }

kmeansErrorStatus kmeansSynchronize(kmeansHandle h) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  if (h->e.mfma) KM_HIP_CHECK(mfma_pair_join(&h->e), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipStreamSynchronize(h->e.stream), KMEANS_ERROR_COMPUTE);
  return KMEANS_SUCCESS;
}

// initialize (kmeans.cpp:822-907): 1 takes centres (n x k, the reference's
// D x K result layout), 2 seeds by k-means++, 3 samples k distinct rows with
// the seed, anything else takes the rows the reference picks.
kmeansErrorStatus kmeansInitialize(kmeansHandle h, int initAlgorithm, uint64_t seed,
                                   const float *centers_dxk) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  Engine &e = h->e;
  if (initAlgorithm == 1) {
    if (!centers_dxk) return KMEANS_ERROR_INVALID_ARGUMENT;
    std::vector<float> c((size_t)e.k * e.n);
    for (int j = 0; j < e.k; ++j)
      for (int q = 0; q < e.n; ++q) c[(size_t)j * e.n + q] = centers_dxk[(size_t)q * e.k + j];
    return e.setCenters(c.data());
  }
  if (initAlgorithm == 2) return e.initialize_pp(seed);
  return e.initialize(seed, initAlgorithm != 3);
}

// closestCenters (kmeans.cpp:911-1014): one assignment of every resident
// point to the current centres (certified re-resolve included for BF16/FP16).
kmeansErrorStatus kmeansAssign(kmeansHandle h) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  Engine &e = h->e;
  if (!e.dX || !e.centers_set) return KMEANS_ERROR_CLOSEST_CENTERS;
  kmeansErrorStatus st = e.closestCenters(0, e.m);
  if (st != KMEANS_SUCCESS) return st;
  if (e.opt.precision != KM_PREC_FP32)
    KM_HIP_CHECK(mfma_reresolve(&e, 0, e.m), KMEANS_ERROR_CLOSEST_CENTERS);
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_CLOSEST_CENTERS);
  return KMEANS_SUCCESS;
}

// clusterCenters (kmeans.cpp:1019-1041): new centres from the current labels.
kmeansErrorStatus kmeansUpdateCenters(kmeansHandle h) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  Engine &e = h->e;
  if (!e.dX || !e.centers_set) return KMEANS_ERROR_CLUSTER_CENTERS;
  kmeansErrorStatus st = e.clusterCenters(0, e.m);
  if (st != KMEANS_SUCCESS) return st;
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_CLUSTER_CENTERS);
  return KMEANS_SUCCESS;
}

// compactness (kmeans.cpp:1045-1104): sum of squared distances of the points
// to the centres of their current labels (FP64, canonical order).
kmeansErrorStatus kmeansCompactness(kmeansHandle h, double *out) {
  if (!h || !out) return KMEANS_ERROR_INVALID_ARGUMENT;
  Engine &e = h->e;
  if (!e.dX || !e.centers_set) return KMEANS_ERROR_COMPACTNESS;
  double *dv = e.ws.compj + e.k;  // the spare slot past the per-centre terms
  KM_HIP_CHECK(launch_compactness(e.stream, e.m, e.n, e.k, e.dX, e.dlabels, e.dC, dv, e.ws),
               KMEANS_ERROR_COMPACTNESS);
  KM_HIP_CHECK(hipMemcpyAsync(out, dv, sizeof(double), hipMemcpyDeviceToHost, e.stream),
               KMEANS_ERROR_COMPACTNESS);
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_COMPACTNESS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansReferenceInitRows(long m, int k, int *rows) {
  if (m <= 0 || m > 0x7fffffffL || k <= 0 || k > m || !rows) return KMEANS_ERROR_INVALID_ARGUMENT;
  km::reference_init_rows(m, k, rows);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansSetMiniBatch(kmeansHandle h, long rows) {
  if (!h || rows < 0 || rows > h->e.m) return KMEANS_ERROR_INVALID_ARGUMENT;
  h->e.mb_rows = rows;
  h->e.mb_index = 0;
  return KMEANS_SUCCESS;
}

static kmeansErrorStatus compute_f(const float *data, int m, int n, int k, double criterion,
                                   int maxIters, int initAlgorithm, int fullGemm, float *result,
                                   double miniBatchFraction) {
  if (!data || !result) return KMEANS_ERROR_INVALID_ARGUMENT;
  kmeansHandle h = nullptr;
  kmeansOptions opt = {};
  opt.full_gemm = fullGemm ? 1 : 0;
  kmeansErrorStatus st = kmeansCreate(&h, m, n, k, &opt);
  if (st != KMEANS_SUCCESS) return st;
  std::vector<float> c((size_t)k * n);
  if ((st = kmeansSetData(h, data, 0)) != KMEANS_SUCCESS) goto done;
  if (initAlgorithm == 1) {
    for (int j = 0; j < k; ++j)
      for (int q = 0; q < n; ++q) c[(size_t)j * n + q] = result[(size_t)q * k + j];
    st = kmeansSetCenters(h, c.data());
  } else if (initAlgorithm == 2) {
    st = h->e.initialize_pp(42);
  } else {
    st = h->e.initialize(42, initAlgorithm != 3);
  }
  if (st != KMEANS_SUCCESS) goto done;
  if (miniBatchFraction < 1.0) {
    long rows = (long)(miniBatchFraction * m);
    if (rows < k) rows = std::min<long>(k, m);
    kmeansSetMiniBatch(h, rows);
  }
  if ((st = kmeansRun(h, maxIters, criterion)) != KMEANS_SUCCESS) goto done;
  if ((st = kmeansGetCenters(h, c.data())) != KMEANS_SUCCESS) goto done;
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < n; ++q) result[(size_t)q * k + j] = c[(size_t)j * n + q];
  st = h->e.getConvergenceFlag();
done:
  kmeansDestroy(h);
  return st;
}

kmeansErrorStatus computeKmeansF(const float *data, int m, int n, int k, float criterion,
                                 int maxIters, int numRetries, int initAlgorithm, int doFullSGEMM,
                                 int matrixFormat, float *result, float miniBatchFraction) {
  (void)numRetries;
  (void)matrixFormat;
  return compute_f(data, m, n, k, criterion, maxIters, initAlgorithm, doFullSGEMM, result,
                   miniBatchFraction);
}

kmeansErrorStatus computeKmeansD(const double *data, int m, int n, int k, double criterion,
                                 int maxIters, int numRetries, int initAlgorithm, int doFullSGEMM,
                                 int matrixFormat, double *result, double miniBatchFraction) {
  (void)numRetries;
  (void)matrixFormat;
  return km::compute_d(data, m, n, k, criterion, maxIters, initAlgorithm, doFullSGEMM, result,
                       miniBatchFraction);
}

// ------------------------------------------------------------ kernel API
#define KM_K_CHECK(expr, code)                                                        \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "kmeans_amd: %s: %s\n", #expr, hipGetErrorString(e_));          \
      return code;                                                                    \
    }                                                                                 \
  } while (0)

kmeansCudaErrorStatus kmColNorms(void *stream, int k, int d, const float *C, float *nc) {
  if (k <= 0 || d <= 0 || !C || !nc) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_col_norms((hipStream_t)stream, k, d, C, nc), ERROR_COLNORMALIZE);
  return NO_ERROR;
}

int kmCentersLd(int k) { return km::centers_ld(k); }

kmeansCudaErrorStatus kmTransposeCenters(void *stream, int k, int d, const float *C, float *Ct) {
  if (k <= 0 || d <= 0 || !C || !Ct) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_transpose((hipStream_t)stream, k, d, C, Ct, km::centers_ld(k)),
             ERROR_CLOSESTCENTERS);
  return NO_ERROR;
}

kmeansCudaErrorStatus kmAssignF32(void *stream, long m, const int *m_dev, int d, const float *X,
                                  const int *rows, int k, const float *Ct, const float *nc,
                                  int *labels, float *best) {
  if (m < 0 || d <= 0 || k <= 0 || !X || !Ct || !nc || !labels) return ERROR_INVALID_ARGUMENT;
  if (m_dev && !rows) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_assign_f32((hipStream_t)stream, m, m_dev, d, X, rows, k, Ct,
                                   km::centers_ld(k), nc, labels, best),
             ERROR_CLOSESTCENTERS);
  return NO_ERROR;
}

size_t kmUpdateWorkspaceSize(long n, int d, int k) {
  return km::update_ws_layout(n, d, k, nullptr, nullptr);
}

kmeansCudaErrorStatus kmUpdate(void *stream, long n, int d, int k, const float *X,
                               const int *labels, const float *C_old, float *C_new, float *Ct_new,
                               float *nc_new, long long *counts, int *empty, void *workspace,
                               size_t workspace_bytes) {
  if (n < 0 || d <= 0 || k <= 0 || !X || !labels || !C_old || !C_new || !workspace)
    return ERROR_INVALID_ARGUMENT;
  km::UpdateWs ws;
  if (km::update_ws_layout(n, d, k, (char *)workspace, &ws) > workspace_bytes)
    return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_update((hipStream_t)stream, n, d, k, X, labels, C_old, C_new,
                               Ct_new, km::centers_ld(k), nc_new, counts, empty, ws),
             ERROR_CLUSTERCENTERS);
  return NO_ERROR;
}

kmeansCudaErrorStatus kmCompactness(void *stream, long n, int d, int k, const float *X,
                                    const int *labels, const float *C, double *out,
                                    void *workspace, size_t workspace_bytes) {
  if (n < 0 || d <= 0 || k <= 0 || !X || !labels || !C || !out || !workspace)
    return ERROR_INVALID_ARGUMENT;
  km::UpdateWs ws;
  if (km::update_ws_layout(n, d, k, (char *)workspace, &ws) > workspace_bytes)
    return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_compactness((hipStream_t)stream, n, d, k, X, labels, C, out, ws),
             ERROR_COMPACTNESS);
  return NO_ERROR;
}

kmeansCudaErrorStatus kmRowNorms(void *stream, long m0, long m, int d, const float *X, float *out) {
  if (m0 < 0 || m < 0 || d <= 0 || !X || !out) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_row_norms((hipStream_t)stream, m, d, X, out + m0), ERROR_ROWNORMALIZE);
  return NO_ERROR;
}

kmeansCudaErrorStatus kmGenerateMixture(void *stream, uint64_t seed, long i0, long n, int d,
                                        int ncomp, const float *means, const float *sigmas,
                                        const double *cdf, float *X, int *z) {
  if (n < 0 || d <= 0 || ncomp <= 0 || !means || !sigmas || !cdf || !X) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_generate((hipStream_t)stream, seed, i0, n, d, ncomp, means, sigmas, cdf, X,
                                 z),
             ERROR_INVALID_ARGUMENT);
  return NO_ERROR;
}

kmeansCudaErrorStatus kmGatherRows(void *stream, long n, int d, const float *X, const int *idx,
                                   float *out) {
  if (n < 0 || d <= 0 || !X || !idx || !out) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::launch_gather_rows((hipStream_t)stream, n, d, X, idx, out), ERROR_INVALID_ARGUMENT);
  return NO_ERROR;
}

size_t kmMlpParamCount(int din, int hidden, int groups) {
  return km::mlp_param_count(km::MlpShape{din, hidden, groups});
}

size_t kmMlpWorkspaceSize(int din, int hidden, int groups, int batch) {
  return km::mlp_ws_bytes(km::MlpShape{din, hidden, groups}, batch);
}

kmeansCudaErrorStatus kmMlpForwardBackward(void *stream, int din, int hidden, int groups,
                                           int batch, const uint16_t *X, const int *targets,
                                           const float *P, float *G, float *loss,
                                           void *workspace, size_t workspace_bytes) {
  if (din <= 0 || din % 16 || hidden <= 0 || hidden % 16 || groups <= 0 || groups % 16 ||
      batch <= 0 || !X ||
      !targets || !P || !G || !workspace)
    return ERROR_INVALID_ARGUMENT;
  const km::MlpShape sh{din, hidden, groups};
  if (km::mlp_ws_bytes(sh, batch) > workspace_bytes) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(km::mlp_fwd_bwd((hipStream_t)stream, sh, batch, X, targets, P, G, loss, workspace),
             ERROR_MATMATMULT);
  return NO_ERROR;
}

}  // extern "C"

// ------------------------------------------------------------ legacy names
// Shared bodies of the ...F and ...D kernel names (KmeansCudaKernels.h:
// 238-474); the element type is that of the points and centres.
namespace km {

template <typename T>
static kmeansCudaErrorStatus legacy_closest(int m0, int nRowsA, int nColsA, const T *A, int nColsB,
                                            const T *B, const T *nc, int *CindicesFinal) {
  if (nRowsA < 0 || nColsA <= 0 || nColsB <= 0 || !A || !B || !nc || !CindicesFinal)
    return ERROR_INVALID_ARGUMENT;
  const int k = nColsB, d = nColsA, ld = centers_ld(k);
  const T *Ct = B;
  std::lock_guard<std::mutex> g(g_legacy_mu);
  // the kernel reads the D x K centres four at a time from rows of ld
  // (a multiple of 256) entries: pad them unless B already has that shape
  if (ld != k || ((uintptr_t)B & (4 * sizeof(T) - 1)) != 0) {
    const size_t need = sizeof(T) * (size_t)ld * d;
    if (need > g_legacy_ct_bytes) {
      hipFree(g_legacy_ct);
      g_legacy_ct = nullptr;
      g_legacy_ct_bytes = 0;
      KM_K_CHECK(hipMalloc(&g_legacy_ct, need), ERROR_CLOSESTCENTERS);
      g_legacy_ct_bytes = need;
    }
    KM_K_CHECK(hipMemset(g_legacy_ct, 0, need), ERROR_CLOSESTCENTERS);
    KM_K_CHECK(hipMemcpy2D(g_legacy_ct, sizeof(T) * ld, B, sizeof(T) * k, sizeof(T) * k, d,
                           hipMemcpyDeviceToDevice),
               ERROR_CLOSESTCENTERS);
    Ct = reinterpret_cast<const T *>(g_legacy_ct);
  }
  hipError_t e;
  if constexpr (sizeof(T) == 4)
    e = launch_assign_f32(nullptr, nRowsA, nullptr, d, A, nullptr, k, Ct, ld, nc,
                          CindicesFinal + m0, nullptr);
  else
    e = launch_assign_f64(nullptr, nRowsA, d, A, k, Ct, ld, nc, CindicesFinal + m0);
  KM_K_CHECK(e, ERROR_CLOSESTCENTERS);
  return NO_ERROR;
}

// centers_large is reused as a k x (n+1) FP64 accumulator and counts_large
// as k int64 counters (the caller zeroes both per iteration, as
// kmeans.cpp:1152-1156 does); on lastTile centers (k x n) and counts are
// written.  Canonical blocks start at each tile's first row.
template <typename T>
static kmeansCudaErrorStatus legacy_cluster(int m0, int m, int n, int k, bool lastTile,
                                            const T *data, const int *indices, T *centers_large,
                                            int *counts_large, T *centers, int *counts) {
  if (m < 0 || n <= 0 || k <= 0 || !data || !indices || !centers_large || !counts_large ||
      !centers)
    return ERROR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(g_legacy_mu);
  UpdateWs ws;
  KM_K_CHECK(legacy_ws(m, n, k, &ws), ERROR_CLUSTERCENTERS);
  double *S = (double *)centers_large;
  long long *cnt = (long long *)counts_large;
  KM_K_CHECK(launch_update_accumulate(nullptr, m, n, k, data, indices + m0, S, cnt, ws),
             ERROR_CLUSTERCENTERS);
  if (lastTile) {
    KM_K_CHECK(launch_update_finalize(nullptr, n, k, S, cnt, (const T *)centers, centers,
                                      (T *)nullptr, 0, (T *)nullptr, ws.cnt, nullptr, nullptr),
               ERROR_CLUSTERCENTERS);
    if (counts) {
      std::vector<long long> c(k);
      std::vector<int> ci(k);
      KM_K_CHECK(hipMemcpy(c.data(), ws.cnt, sizeof(long long) * k, hipMemcpyDeviceToHost),
                 ERROR_CLUSTERCENTERS);
      for (int j = 0; j < k; ++j) ci[j] = (int)c[j];
      KM_K_CHECK(hipMemcpy(counts, ci.data(), sizeof(int) * k, hipMemcpyHostToDevice),
                 ERROR_CLUSTERCENTERS);
    }
  }
  KM_K_CHECK(hipDeviceSynchronize(), ERROR_CLUSTERCENTERS);
  return NO_ERROR;
}

// The 8 bytes after compactness[0] hold the FP64 running total across tiles
// (compactness[2..3] for float, compactness[1] for double); on lastTile
// compactness[0] = (T)total.  The caller zeroes the buffer per iteration.
template <typename T>
static kmeansCudaErrorStatus legacy_compactness(int m0, int m, bool lastTile, int n, int k,
                                                const T *data, const int *indices,
                                                const T *centers, T *compactness) {
  if (m < 0 || n <= 0 || k <= 0 || !data || !indices || !centers || !compactness)
    return ERROR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(g_legacy_mu);
  UpdateWs ws;
  KM_K_CHECK(legacy_ws(m, n, k, &ws), ERROR_COMPACTNESS);
  double *acc = (double *)(compactness + 8 / sizeof(T));
  double *tile = ws.compj + k;  // one spare slot past the per-centre terms
  KM_K_CHECK(launch_compactness(nullptr, m, n, k, data, indices + m0, centers, tile, ws),
             ERROR_COMPACTNESS);
  double tv = 0.0, tot = 0.0;
  KM_K_CHECK(hipMemcpy(&tv, tile, sizeof(double), hipMemcpyDeviceToHost), ERROR_COMPACTNESS);
  KM_K_CHECK(hipMemcpy(&tot, acc, sizeof(double), hipMemcpyDeviceToHost), ERROR_COMPACTNESS);
  tot += tv;
  KM_K_CHECK(hipMemcpy(acc, &tot, sizeof(double), hipMemcpyHostToDevice), ERROR_COMPACTNESS);
  if (lastTile) {
    const T f = (T)tot;
    KM_K_CHECK(hipMemcpy(compactness, &f, sizeof(T), hipMemcpyHostToDevice), ERROR_COMPACTNESS);
  }
  KM_K_CHECK(hipDeviceSynchronize(), ERROR_COMPACTNESS);
  return NO_ERROR;
}

template <typename T>
static kmeansCudaErrorStatus legacy_row_norms(int m0, int m, int n, const T *in, T *out) {
  if (m < 0 || n <= 0 || !in || !out) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(launch_row_norms(nullptr, m, n, in, out + m0), ERROR_ROWNORMALIZE);
  return NO_ERROR;
}

// in is m x n = D x K
template <typename T>
static kmeansCudaErrorStatus legacy_col_norms(int m, int n, const T *in, T *out) {
  if (m <= 0 || n <= 0 || !in || !out) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(launch_col_norms_dk(nullptr, n, m, n, in, out), ERROR_COLNORMALIZE);
  return NO_ERROR;
}

template <typename T>
static kmeansCudaErrorStatus legacy_row_transform_min(int m0, int m, int n, const T *rn,
                                                      const T *cn, const T *input, int *output) {
  if (m0 < 0 || m < 0 || n <= 0 || !rn || !cn || !input || !output) return ERROR_INVALID_ARGUMENT;
  KM_K_CHECK(launch_row_transform_min(nullptr, m0, m, n, rn, cn, input, output, -1),
             ERROR_ROWTRANSFORMMINIMUM);
  return NO_ERROR;
}

}  // namespace km

extern "C" {

// ClosestCentersF/D (KmeansCudaKernels.h:238-273): A = nRowsA x nColsA tile,
// B = nColsA x nColsB centres (D x K), labels to CindicesFinal[m0 + i].
kmeansCudaErrorStatus ClosestCentersF(const int m0, const int nRowsA, const int nColsA,
                                      const float *A, const int nColsB, const float *B,
                                      float *Bpadded, const float *normRowsOfA_squared,
                                      const float *normColsOfB_squared, const int nColsC,
                                      float *C, int *Cindices, int *CindicesFinal) {
  (void)Bpadded; (void)normRowsOfA_squared; (void)nColsC; (void)C; (void)Cindices;
  return km::legacy_closest(m0, nRowsA, nColsA, A, nColsB, B, normColsOfB_squared, CindicesFinal);
}

kmeansCudaErrorStatus ClosestCentersD(const int m0, const int nRowsA, const int nColsA,
                                      const double *A, const int nColsB, const double *B,
                                      double *Bpadded, const double *normRowsOfA_squared,
                                      const double *normColsOfB_squared, const int nColsC,
                                      double *C, int *Cindices, int *CindicesFinal) {
  (void)Bpadded; (void)normRowsOfA_squared; (void)nColsC; (void)C; (void)Cindices;
  return km::legacy_closest(m0, nRowsA, nColsA, A, nColsB, B, normColsOfB_squared, CindicesFinal);
}

// rowTransformMinimumF/D (KmeansCudaKernels.h:290-313)
kmeansCudaErrorStatus rowTransformMinimumF(const int m0, const int m, const int n,
                                           const float *normRowsOfA_squared,
                                           const float *normColsOfB_squared, const float *input,
                                           int *output) {
  return km::legacy_row_transform_min(m0, m, n, normRowsOfA_squared, normColsOfB_squared, input,
                                      output);
}

kmeansCudaErrorStatus rowTransformMinimumD(const int m0, const int m, const int n,
                                           const double *normRowsOfA_squared,
                                           const double *normColsOfB_squared, const double *input,
                                           int *output) {
  return km::legacy_row_transform_min(m0, m, n, normRowsOfA_squared, normColsOfB_squared, input,
                                      output);
}

// ClusterCentersF/D (KmeansCudaKernels.h:337-369)
kmeansCudaErrorStatus ClusterCentersF(const int m0, const int m, const int n, const int k,
                                      const bool lastTile, const float *data, const int *indices,
                                      float *centers_large, int *counts_large, float *centers,
                                      int *counts) {
  return km::legacy_cluster(m0, m, n, k, lastTile, data, indices, centers_large, counts_large,
                            centers, counts);
}

kmeansCudaErrorStatus ClusterCentersD(const int m0, const int m, const int n, const int k,
                                      const bool lastTile, const double *data, const int *indices,
                                      double *centers_large, int *counts_large, double *centers,
                                      int *counts) {
  return km::legacy_cluster(m0, m, n, k, lastTile, data, indices, centers_large, counts_large,
                            centers, counts);
}

// CompactnessF/D (KmeansCudaKernels.h:389-414)
kmeansCudaErrorStatus CompactnessF(const int m0, const int m, const bool lastTile, const int n,
                                   const int k, const float *data, const int *indices,
                                   const float *centers, float *compactness) {
  return km::legacy_compactness(m0, m, lastTile, n, k, data, indices, centers, compactness);
}

kmeansCudaErrorStatus CompactnessD(const int m0, const int m, const bool lastTile, const int n,
                                   const int k, const double *data, const int *indices,
                                   const double *centers, double *compactness) {
  return km::legacy_compactness(m0, m, lastTile, n, k, data, indices, centers, compactness);
}

// rowNormalizeF/D (KmeansCudaKernels.h:428-445), colNormalizeF/D (:458-474)
kmeansCudaErrorStatus rowNormalizeF(const int m0, const int m, const int n, const float *in,
                                    float *out) {
  return km::legacy_row_norms(m0, m, n, in, out);
}

kmeansCudaErrorStatus rowNormalizeD(const int m0, const int m, const int n, const double *in,
                                    double *out) {
  return km::legacy_row_norms(m0, m, n, in, out);
}

kmeansCudaErrorStatus colNormalizeF(const int m, const int n, const float *in, float *out) {
  return km::legacy_col_norms(m, n, in, out);
}

kmeansCudaErrorStatus colNormalizeD(const int m, const int n, const double *in, double *out) {
  return km::legacy_col_norms(m, n, in, out);
}

}  // extern "C"
