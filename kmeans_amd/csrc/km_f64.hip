// Double-precision engine behind computeKmeansD (main/cpp/kmeans.cpp:
// 1364-1409, kmeans.hpp:125-130).
//
// The reference's fused FP64 path never produces labels: its argmax
// reduction for double is an empty stub (main/cuda/reduction.hcu:8-11), so
// every point gets label 0 and the centres become NaN.  This engine runs the
// same Lloyd loop as the FP32 one with double points and centres:
//   assign   k_assign_f64 (fma chain in double from -0.5*nc_j over d
//            ascending, lowest index among the maxima);
//   update   the FP32 path's histogram / scan / stable scatter / chunk sums
//            on double rows (sums were FP64 already), centres S/n in double;
//   compactness  from the per-centre moments, as in the FP32 engine.
// Control flow follows computeStandard (kmeans.cpp:1186-1247) and
// computeMiniBatch (:1253-1312) exactly as the FP32 engine does.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "km_common.hpp"
#include "km_gemm.hpp"

namespace km {

namespace {

struct EngineD {
  long m = 0;
  int n = 0, k = 0, ldc = 0;
  hipStream_t stream = nullptr;
  double *dX = nullptr, *dC = nullptr, *dCt = nullptr, *dnc = nullptr, *dq = nullptr;
  double *dcomp = nullptr;
  double *dxn = nullptr;  // row norms for the library-GEMM path
  int *dlabels = nullptr, *dempty = nullptr, *didx = nullptr;
  long long *dcounts = nullptr;
  char *dws = nullptr;
  UpdateWs ws{};
  GemmAssign *gemm = nullptr;  // doFullSGEMM = 1: rocBLAS DGEMM + row argmin

  ~EngineD() {
    if (stream) hipStreamSynchronize(stream);
    gemm_destroy(gemm);
    hipFree(dX);
    hipFree(dC);
    hipFree(dCt);
    hipFree(dnc);
    hipFree(dq);
    hipFree(dxn);
    hipFree(dcomp);
    hipFree(dlabels);
    hipFree(dempty);
    hipFree(didx);
    hipFree(dcounts);
    hipFree(dws);
    if (stream) hipStreamDestroy(stream);
  }
};

}  // namespace

kmeansErrorStatus compute_d(const double *data, int m, int n, int k, double criterion,
                            int maxIters, int initAlgorithm, int fullGemm, double *result,
                            double miniBatchFraction) {
  if (!data || !result || m <= 0 || n <= 0 || k <= 0 || k > m || k > 16384)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    fprintf(stderr, "kmeans_amd: no HIP device available\n");
    return KMEANS_ERROR_BUILDDATA;
  }
  EngineD E;
  E.m = m;
  E.n = n;
  E.k = k;
  E.ldc = centers_ld(k);
  KM_HIP_CHECK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking), KMEANS_ERROR_BUILDDATA);
  hipStream_t s = E.stream;
  const size_t xb = sizeof(double) * (size_t)m * n;
  KM_HIP_CHECK(hipMalloc(&E.dX, xb), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dC, sizeof(double) * (size_t)k * n), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dCt, sizeof(double) * (size_t)E.ldc * n), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dnc, sizeof(double) * k), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dq, sizeof(double) * (size_t)m), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dlabels, sizeof(int) * (size_t)m), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dempty, sizeof(int)), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dcounts, sizeof(long long) * k), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMalloc(&E.dcomp, sizeof(double) * std::max(maxIters, 1)),
               KMEANS_ERROR_BUILDDATA);
  const size_t wsb = update_ws_layout(m, n, k, nullptr, nullptr);
  KM_HIP_CHECK(hipMalloc(&E.dws, wsb), KMEANS_ERROR_BUILDDATA);
  update_ws_layout(m, n, k, E.dws, &E.ws);
  // buildData (kmeans.cpp:540-695, without the row shuffle): points and
  // their FP64 squared norms for the compactness moments
  KM_HIP_CHECK(hipMemcpyAsync(E.dX, data, xb, hipMemcpyHostToDevice, s), KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(hipMemsetAsync(E.dCt, 0, sizeof(double) * (size_t)E.ldc * n, s),
               KMEANS_ERROR_BUILDDATA);
  KM_HIP_CHECK(launch_point_sq(s, m, n, E.dX, E.dq), KMEANS_ERROR_BUILDDATA);

  // initialize (kmeans.cpp:822-907): result (D x K) or k sampled rows
  if (initAlgorithm == 1) {
    std::vector<double> c((size_t)k * n);
    for (int j = 0; j < k; ++j)
      for (int q = 0; q < n; ++q) c[(size_t)j * n + q] = result[(size_t)q * k + j];
    KM_HIP_CHECK(hipMemcpyAsync(E.dC, c.data(), sizeof(double) * c.size(), hipMemcpyHostToDevice,
                                s),
                 KMEANS_ERROR_INITIALIZE);
    KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_INITIALIZE);
  } else {
    KM_HIP_CHECK(hipMalloc(&E.didx, sizeof(int) * k), KMEANS_ERROR_INITIALIZE);
    if (initAlgorithm == 2) {  // k-means++ seeding
      KM_HIP_CHECK(kmeanspp(s, m, n, k, E.dX, 42, E.didx), KMEANS_ERROR_INITIALIZE);
    } else {
      std::vector<int> idx(k);
      sample_rows(42, m, k, idx.data());
      KM_HIP_CHECK(hipMemcpyAsync(E.didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, s),
                   KMEANS_ERROR_INITIALIZE);
      KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_INITIALIZE);
    }
    KM_HIP_CHECK(launch_gather_rows(s, k, n, E.dX, E.didx, E.dC), KMEANS_ERROR_INITIALIZE);
    KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_INITIALIZE);
  }
  KM_HIP_CHECK(launch_transpose(s, k, n, E.dC, E.dCt, E.ldc), KMEANS_ERROR_TRANSPOSE);
  KM_HIP_CHECK(launch_col_norms(s, k, n, E.dC, E.dnc), KMEANS_ERROR_TRANSPOSE);

  // mini-batch windows [0,mb), [mb,2mb), ..., then back to row 0
  long mb = 0;
  if (miniBatchFraction < 1.0) {
    mb = (long)(miniBatchFraction * m);
    if (mb < k) mb = std::min<long>(k, m);
  }
  if (fullGemm) {  // the data's row norms, once (k_row_norms, as per tile)
    KM_HIP_CHECK(hipMalloc(&E.dxn, sizeof(double) * (size_t)m), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(launch_row_norms(s, m, n, E.dX, E.dxn), KMEANS_ERROR_BUILDDATA);
  }
  const bool sync_each = criterion >= 0.0;
  double prev = 0.0, relErr = 1.0;
  int it = 0;
  long mb_index = 0;
  while (it < maxIters && relErr > criterion) {
    long r0 = 0, rows = m;
    if (mb > 0) {
      const long per_epoch = (m + mb - 1) / mb;
      r0 = (mb_index++ % per_epoch) * mb;
      rows = std::min(mb, (long)m - r0);
    }
    if (fullGemm)
      KM_HIP_CHECK(gemm_assign(&E.gemm, s, E.dX + r0 * n, E.dxn + r0, rows, n, k, E.dCt, E.ldc,
                               E.dnc, E.dlabels + r0),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    else
      KM_HIP_CHECK(launch_assign_f64(s, rows, n, E.dX + r0 * n, k, E.dCt, E.ldc, E.dnc,
                                     E.dlabels + r0),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    KM_HIP_CHECK(launch_update(s, rows, n, k, E.dX + r0 * n, E.dlabels + r0, E.dq + r0, E.dC, E.dC,
                               E.dCt, E.ldc, E.dnc, E.dcounts, E.dempty, E.ws),
                 KMEANS_ERROR_CLUSTER_CENTERS);
    KM_HIP_CHECK(launch_comp_reduce(s, k, E.ws.compj, E.dcomp + it), KMEANS_ERROR_COMPACTNESS);
    ++it;
    if (sync_each) {
      double comp = 0.0;
      KM_HIP_CHECK(hipMemcpyAsync(&comp, E.dcomp + it - 1, sizeof(double), hipMemcpyDeviceToHost,
                                  s),
                   KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_COMPUTE);
      relErr = fabs(comp - prev) / fabs(comp);
      prev = comp;
      if (mb > 0 && relErr <= criterion) break;
    } else {
      relErr = 1.0 + fabs(criterion);  // fixed iteration count, no host round trip
    }
  }
  if (!sync_each && it > 0) {
    std::vector<double> hist(it);
    KM_HIP_CHECK(hipMemcpyAsync(hist.data(), E.dcomp, sizeof(double) * it, hipMemcpyDeviceToHost,
                                s),
                 KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_COMPUTE);
    for (int t = 0; t < it; ++t) {
      relErr = fabs(hist[t] - prev) / fabs(hist[t]);
      prev = hist[t];
    }
  }
  std::vector<double> c((size_t)k * n);
  KM_HIP_CHECK(hipMemcpyAsync(c.data(), E.dC, sizeof(double) * c.size(), hipMemcpyDeviceToHost, s),
               KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_COMPUTE);
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < n; ++q) result[(size_t)q * k + j] = c[(size_t)j * n + q];
  return relErr <= criterion ? KMEANS_SUCCESS : KMEANS_NONCONVERGENCE;
}

}  // namespace km
