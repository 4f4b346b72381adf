// Double-precision engine behind computeKmeansD (main/cpp/kmeans.cpp:
// 1364-1409, kmeans.hpp:125-130) and the kmeans<double> class's steps.
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
#include <new>
#include <vector>

#include "km_common.hpp"
#include "km_gemm.hpp"

namespace km {

struct EngineD {
  long m = 0;
  int n = 0, k = 0, ldc = 0;
  bool full_gemm = false;
  hipStream_t stream = nullptr;
  double *dX = nullptr, *dC = nullptr, *dCt = nullptr, *dnc = nullptr;
  const double *dXb = nullptr;  // the points (dX, or a borrowed device array)
  double *dcomp = nullptr;
  int comp_slots = 0;
  double *dxn = nullptr;  // row norms for the library-GEMM path
  int *dlabels = nullptr, *dempty = nullptr, *didx = nullptr;
  long long *dcounts = nullptr;
  char *dws = nullptr;
  UpdateWs ws{};
  GemmAssign *gemm = nullptr;  // doFullSGEMM = 1: rocBLAS DGEMM + row argmin
  bool centers_set = false;
  long mb = 0, mb_index = 0;  // mini-batch rows (0: standard loop) and window
  int iters = 0;
  double prev = 0.0, relErr = 1.0, criterion = -1.0;
  kmeansStats stats{};

  ~EngineD() {
    if (stream) hipStreamSynchronize(stream);
    gemm_destroy(gemm);
    hipFree(dX);
    hipFree(dC);
    hipFree(dCt);
    hipFree(dnc);
    hipFree(dxn);
    hipFree(dcomp);
    hipFree(dlabels);
    hipFree(dempty);
    hipFree(didx);
    hipFree(dcounts);
    hipFree(dws);
    if (stream) hipStreamDestroy(stream);
  }

  kmeansErrorStatus create(long m_, int n_, int k_, const kmeansOptions *o) {
    if (m_ <= 0 || n_ <= 0 || k_ <= 0 || k_ > m_ || m_ > 0x7fffffffL || k_ > 16384)
      return KMEANS_ERROR_INVALID_ARGUMENT;
    if (o && o->precision != KM_PREC_FP32) return KMEANS_ERROR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
      fprintf(stderr, "kmeans_amd: no HIP device available\n");
      return KMEANS_ERROR_BUILDDATA;
    }
    m = m_;
    n = n_;
    k = k_;
    ldc = centers_ld(k);
    full_gemm = o && o->full_gemm;
    if (o) KM_HIP_CHECK(hipSetDevice(o->device), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dC, sizeof(double) * (size_t)k * n), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dCt, sizeof(double) * (size_t)ldc * n), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMemsetAsync(dCt, 0, sizeof(double) * (size_t)ldc * n, stream),
                 KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dnc, sizeof(double) * k), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dlabels, sizeof(int) * (size_t)m), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMemsetAsync(dlabels, 0, sizeof(int) * (size_t)m, stream),
                 KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dempty, sizeof(int)), KMEANS_ERROR_BUILDDATA);
    KM_HIP_CHECK(hipMalloc(&dcounts, sizeof(long long) * k), KMEANS_ERROR_BUILDDATA);
    const size_t wsb = update_ws_layout(m, n, k, nullptr, nullptr);
    KM_HIP_CHECK(hipMalloc(&dws, wsb), KMEANS_ERROR_BUILDDATA);
    update_ws_layout(m, n, k, dws, &ws);
    return KMEANS_SUCCESS;
  }

  // buildData (kmeans.cpp:540-695, without the row shuffle): the points
  kmeansErrorStatus setData(const double *data, int on_device) {
    if (!data) return KMEANS_ERROR_INVALID_ARGUMENT;
    if (on_device) {
      dXb = data;
    } else {
      const size_t xb = sizeof(double) * (size_t)m * n;
      if (!dX) KM_HIP_CHECK(hipMalloc(&dX, xb), KMEANS_ERROR_BUILDDATA);
      KM_HIP_CHECK(hipMemcpyAsync(dX, data, xb, hipMemcpyHostToDevice, stream),
                   KMEANS_ERROR_BUILDDATA);
      dXb = dX;
    }
    if (full_gemm) {  // the data's row norms, once (k_row_norms, as per tile)
      if (!dxn) KM_HIP_CHECK(hipMalloc(&dxn, sizeof(double) * (size_t)m), KMEANS_ERROR_BUILDDATA);
      KM_HIP_CHECK(launch_row_norms(stream, m, n, dXb, dxn), KMEANS_ERROR_BUILDDATA);
    }
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_BUILDDATA);
    return KMEANS_SUCCESS;
  }

  kmeansErrorStatus transpose() {
    KM_HIP_CHECK(launch_transpose(stream, k, n, dC, dCt, ldc), KMEANS_ERROR_TRANSPOSE);
    KM_HIP_CHECK(launch_col_norms(stream, k, n, dC, dnc), KMEANS_ERROR_TRANSPOSE);
    centers_set = true;
    relErr = 1.0;
    prev = 0.0;
    return KMEANS_SUCCESS;
  }

  // centres k x n host
  kmeansErrorStatus setCenters(const double *c) {
    if (!c) return KMEANS_ERROR_INVALID_ARGUMENT;
    KM_HIP_CHECK(hipMemcpyAsync(dC, c, sizeof(double) * (size_t)k * n, hipMemcpyHostToDevice,
                                stream),
                 KMEANS_ERROR_INITIALIZE);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_INITIALIZE);
    return transpose();
  }

  // initialize (kmeans.cpp:822-907): 1 takes centres_dxk (n x k, the
  // reference's result layout), 2 seeds by k-means++, 3 samples k distinct
  // rows with the seed, anything else takes the rows the reference picks
  kmeansErrorStatus initialize(int initAlgorithm, uint64_t seed, const double *centers_dxk) {
    if (!dXb) return KMEANS_ERROR_INITIALIZE;
    if (initAlgorithm == 1) {
      if (!centers_dxk) return KMEANS_ERROR_INVALID_ARGUMENT;
      std::vector<double> c((size_t)k * n);
      for (int j = 0; j < k; ++j)
        for (int q = 0; q < n; ++q) c[(size_t)j * n + q] = centers_dxk[(size_t)q * k + j];
      return setCenters(c.data());
    }
    if (!didx) KM_HIP_CHECK(hipMalloc(&didx, sizeof(int) * k), KMEANS_ERROR_INITIALIZE);
    if (initAlgorithm == 2) {
      KM_HIP_CHECK(kmeanspp(stream, m, n, k, dXb, seed, didx), KMEANS_ERROR_INITIALIZE);
    } else {
      std::vector<int> idx(k);
      if (initAlgorithm == 3)
        sample_rows(seed, m, k, idx.data());
      else
        reference_init_rows(m, k, idx.data());
      KM_HIP_CHECK(hipMemcpyAsync(didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice,
                                  stream),
                   KMEANS_ERROR_INITIALIZE);
      KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_INITIALIZE);
    }
    KM_HIP_CHECK(launch_gather_rows(stream, k, n, dXb, didx, dC), KMEANS_ERROR_INITIALIZE);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_INITIALIZE);
    return transpose();
  }

  kmeansErrorStatus closestCenters(long r0, long rows) {
    if (full_gemm)
      KM_HIP_CHECK(gemm_assign(&gemm, stream, dXb + r0 * n, dxn + r0, rows, n, k, dCt, ldc, dnc,
                               dlabels + r0),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    else
      KM_HIP_CHECK(launch_assign_f64(stream, rows, n, dXb + r0 * n, k, dCt, ldc, dnc,
                                     dlabels + r0),
                   KMEANS_ERROR_CLOSEST_CENTERS);
    return KMEANS_SUCCESS;
  }

  kmeansErrorStatus clusterCenters(long r0, long rows) {
    KM_HIP_CHECK(launch_update(stream, rows, n, k, dXb + r0 * n, dlabels + r0, dC, dC,
                               dCt, ldc, dnc, dcounts, dempty, ws),
                 KMEANS_ERROR_CLUSTER_CENTERS);
    return KMEANS_SUCCESS;
  }

  kmeansErrorStatus ensure_slots(int nslots) {
    if (nslots <= comp_slots) return KMEANS_SUCCESS;
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_BUILDDATA);
    hipFree(dcomp);
    dcomp = nullptr;
    comp_slots = 0;
    KM_HIP_CHECK(hipMalloc(&dcomp, sizeof(double) * nslots), KMEANS_ERROR_BUILDDATA);
    comp_slots = nslots;
    return KMEANS_SUCCESS;
  }

  // while (it < maxIters && relErr > criterion), relErr = |comp - prev| /
  // |comp| (kmeans.cpp:1194,1235-1237); mini-batch windows [0,mb), [mb,2mb),
  // ..., then back to row 0, with an early return once relErr <= criterion
  // (:1253-1312).  criterion < 0: no host round trip inside the loop.
  kmeansErrorStatus run(int maxIt, double crit) {
    if (!dXb || !centers_set) return KMEANS_ERROR_COMPUTE;
    criterion = crit;
    if (relErr <= criterion) relErr = 1.0 + fabs(criterion);
    kmeansErrorStatus err;
    if ((err = ensure_slots(std::max(maxIt, 1))) != KMEANS_SUCCESS) return err;
    const bool sync_each = criterion >= 0.0;
    const int it_end = iters + maxIt;
    int slot = 0;
    while (iters < it_end && relErr > criterion) {
      long r0 = 0, rows = m;
      if (mb > 0) {
        const long per_epoch = (m + mb - 1) / mb;
        r0 = (mb_index++ % per_epoch) * mb;
        rows = std::min(mb, m - r0);
      }
      if ((err = closestCenters(r0, rows)) != KMEANS_SUCCESS) return err;
      if ((err = clusterCenters(r0, rows)) != KMEANS_SUCCESS) return err;
      KM_HIP_CHECK(launch_comp_reduce(stream, k, ws.compj, dcomp + slot), KMEANS_ERROR_COMPACTNESS);
      ++iters;
      ++slot;
      if (sync_each) {
        double comp = 0.0;
        KM_HIP_CHECK(hipMemcpyAsync(&comp, dcomp + slot - 1, sizeof(double),
                                    hipMemcpyDeviceToHost, stream),
                     KMEANS_ERROR_COMPUTE);
        KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_COMPUTE);
        relErr = fabs(comp - prev) / fabs(comp);
        prev = comp;
        if (mb > 0 && relErr <= criterion) break;
      } else {
        relErr = 1.0 + fabs(criterion);
      }
    }
    if (!sync_each && slot > 0) {
      std::vector<double> hist(slot);
      KM_HIP_CHECK(hipMemcpyAsync(hist.data(), dcomp, sizeof(double) * slot,
                                  hipMemcpyDeviceToHost, stream),
                   KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_COMPUTE);
      for (int t = 0; t < slot; ++t) {
        relErr = fabs(hist[t] - prev) / fabs(hist[t]);
        prev = hist[t];
      }
    }
    stats.iterations = iters;
    stats.compactness = prev;
    stats.rel_err = relErr;
    return KMEANS_SUCCESS;
  }

  kmeansErrorStatus getCenters(double *out) {
    KM_HIP_CHECK(hipMemcpyAsync(out, dC, sizeof(double) * (size_t)k * n, hipMemcpyDeviceToHost,
                                stream),
                 KMEANS_ERROR_GET_CENTERS);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_GET_CENTERS);
    return KMEANS_SUCCESS;
  }

  kmeansErrorStatus getLabels(int *out) {
    KM_HIP_CHECK(hipMemcpyAsync(out, dlabels, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost,
                                stream),
                 KMEANS_ERROR_GET_CENTERS);
    KM_HIP_CHECK(hipStreamSynchronize(stream), KMEANS_ERROR_GET_CENTERS);
    return KMEANS_SUCCESS;
  }
};

kmeansErrorStatus compute_d(const double *data, int m, int n, int k, double criterion,
                            int maxIters, int initAlgorithm, int fullGemm, double *result,
                            double miniBatchFraction) {
  if (!data || !result) return KMEANS_ERROR_INVALID_ARGUMENT;
  EngineD E;
  kmeansOptions opt = {};
  opt.full_gemm = fullGemm ? 1 : 0;
  kmeansErrorStatus st;
  if ((st = E.create(m, n, k, &opt)) != KMEANS_SUCCESS) return st;
  if ((st = E.setData(data, 0)) != KMEANS_SUCCESS) return st;
  if ((st = E.initialize(initAlgorithm, 42, result)) != KMEANS_SUCCESS)
    return st;
  if (miniBatchFraction < 1.0) {
    E.mb = (long)(miniBatchFraction * m);
    if (E.mb < k) E.mb = std::min<long>(k, m);
  }
  if ((st = E.run(maxIters, criterion)) != KMEANS_SUCCESS) return st;
  std::vector<double> c((size_t)k * n);
  if ((st = E.getCenters(c.data())) != KMEANS_SUCCESS) return st;
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < n; ++q) result[(size_t)q * k + j] = c[(size_t)j * n + q];
  return E.relErr <= criterion ? KMEANS_SUCCESS : KMEANS_NONCONVERGENCE;
}

}  // namespace km

struct kmeansEngineD {
  km::EngineD e;
};

extern "C" {

kmeansErrorStatus kmeansCreateD(kmeansHandleD *h, long m, int n, int k, const kmeansOptions *opt) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  *h = nullptr;
  kmeansEngineD *p = new (std::nothrow) kmeansEngineD();
  if (!p) return KMEANS_ERROR_BUILDDATA;
  kmeansErrorStatus st = p->e.create(m, n, k, opt);
  if (st != KMEANS_SUCCESS) {
    delete p;
    return st;
  }
  *h = p;
  return KMEANS_SUCCESS;
}

void kmeansDestroyD(kmeansHandleD h) { delete h; }

kmeansErrorStatus kmeansSetDataD(kmeansHandleD h, const double *data, int on_device) {
  return h ? h->e.setData(data, on_device) : KMEANS_ERROR_INVALID_ARGUMENT;
}

kmeansErrorStatus kmeansSetCentersD(kmeansHandleD h, const double *centers) {
  return h ? h->e.setCenters(centers) : KMEANS_ERROR_INVALID_ARGUMENT;
}

kmeansErrorStatus kmeansInitializeD(kmeansHandleD h, int initAlgorithm, uint64_t seed,
                                    const double *centers_dxk) {
  return h ? h->e.initialize(initAlgorithm, seed, centers_dxk) : KMEANS_ERROR_INVALID_ARGUMENT;
}

kmeansErrorStatus kmeansAssignD(kmeansHandleD h) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  km::EngineD &e = h->e;
  if (!e.dXb || !e.centers_set) return KMEANS_ERROR_CLOSEST_CENTERS;
  kmeansErrorStatus st = e.closestCenters(0, e.m);
  if (st != KMEANS_SUCCESS) return st;
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_CLOSEST_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansUpdateCentersD(kmeansHandleD h) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  km::EngineD &e = h->e;
  if (!e.dXb || !e.centers_set) return KMEANS_ERROR_CLUSTER_CENTERS;
  kmeansErrorStatus st = e.clusterCenters(0, e.m);
  if (st != KMEANS_SUCCESS) return st;
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_CLUSTER_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansCompactnessD(kmeansHandleD h, double *out) {
  if (!h || !out) return KMEANS_ERROR_INVALID_ARGUMENT;
  km::EngineD &e = h->e;
  if (!e.dXb || !e.centers_set) return KMEANS_ERROR_COMPACTNESS;
  double *dv = e.ws.compj + e.k;  // the spare slot past the per-centre terms
  KM_HIP_CHECK(km::launch_compactness(e.stream, e.m, e.n, e.k, e.dXb, e.dlabels, e.dC, dv,
                                      e.ws),
               KMEANS_ERROR_COMPACTNESS);
  KM_HIP_CHECK(hipMemcpyAsync(out, dv, sizeof(double), hipMemcpyDeviceToHost, e.stream),
               KMEANS_ERROR_COMPACTNESS);
  KM_HIP_CHECK(hipStreamSynchronize(e.stream), KMEANS_ERROR_COMPACTNESS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansRunD(kmeansHandleD h, int maxIters, double criterion) {
  if (!h || maxIters < 0) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.run(maxIters, criterion);
}

kmeansErrorStatus kmeansSetMiniBatchD(kmeansHandleD h, long rows) {
  if (!h || rows < 0 || rows > h->e.m) return KMEANS_ERROR_INVALID_ARGUMENT;
  h->e.mb = rows;
  h->e.mb_index = 0;
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansGetCentersD(kmeansHandleD h, double *centers) {
  if (!h || !centers) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.getCenters(centers);
}

kmeansErrorStatus kmeansGetLabelsD(kmeansHandleD h, int *labels) {
  if (!h || !labels) return KMEANS_ERROR_INVALID_ARGUMENT;
  return h->e.getLabels(labels);
}

kmeansErrorStatus kmeansGetStatsD(kmeansHandleD h, kmeansStats *stats) {
  if (!h || !stats) return KMEANS_ERROR_INVALID_ARGUMENT;
  *stats = h->e.stats;
  return KMEANS_SUCCESS;
}

}  // extern "C"
