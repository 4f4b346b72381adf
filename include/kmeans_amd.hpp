/*
 * kmeans_amd.hpp -- the reference's host class kmeans<TYPE>
 * (main/cpp/kmeans.hpp:174-312) over the C-ABI of kmeans_amd.h, header only.
 *
 * Same constructor arguments and public method names, so code written
 * against the reference's class compiles against this one after the include
 * and namespace change.  kmeans<float> and kmeans<double> drive an engine
 * handle (kmeansCreate / kmeansCreateD) step by step: closestCenters,
 * clusterCenters and compactness are one C call each.  Methods whose work
 * the engine does implicitly (transpose, reset, computeTiling,
 * copyTileToDevice, constructMiniBatch) return KMEANS_SUCCESS.  Added beyond
 * the reference: getLabels, getCompactness, setInitialCenters.
 */
#ifndef KMEANS_AMD_HPP_
#define KMEANS_AMD_HPP_

#include <stdint.h>

#include <vector>

#include "kmeans_amd.h"

namespace kmeans_amd {

namespace detail {

// The engine calls of one element type.
template <class T>
struct Api;

template <>
struct Api<float> {
  typedef kmeansHandle H;
  static kmeansErrorStatus create(H *h, long m, int n, int k, const kmeansOptions *o) {
    return kmeansCreate(h, m, n, k, o);
  }
  static void destroy(H h) { kmeansDestroy(h); }
  static kmeansErrorStatus setData(H h, const float *x) { return kmeansSetData(h, x, 0); }
  static kmeansErrorStatus setCenters(H h, const float *c) { return kmeansSetCenters(h, c); }
  static kmeansErrorStatus initialize(H h, int alg, uint64_t seed, const float *c) {
    return kmeansInitialize(h, alg, seed, c);
  }
  static kmeansErrorStatus assign(H h) { return kmeansAssign(h); }
  static kmeansErrorStatus update(H h) { return kmeansUpdateCenters(h); }
  static kmeansErrorStatus compactness(H h, double *o) { return kmeansCompactness(h, o); }
  static kmeansErrorStatus run(H h, int it, double c) { return kmeansRun(h, it, c); }
  static kmeansErrorStatus miniBatch(H h, long rows) { return kmeansSetMiniBatch(h, rows); }
  static kmeansErrorStatus centers(H h, float *c) { return kmeansGetCenters(h, c); }
  static kmeansErrorStatus labels(H h, int *l) { return kmeansGetLabels(h, l); }
  static kmeansErrorStatus stats(H h, kmeansStats *s) { return kmeansGetStats(h, s); }
};

template <>
struct Api<double> {
  typedef kmeansHandleD H;
  static kmeansErrorStatus create(H *h, long m, int n, int k, const kmeansOptions *o) {
    return kmeansCreateD(h, m, n, k, o);
  }
  static void destroy(H h) { kmeansDestroyD(h); }
  static kmeansErrorStatus setData(H h, const double *x) { return kmeansSetDataD(h, x, 0); }
  static kmeansErrorStatus setCenters(H h, const double *c) { return kmeansSetCentersD(h, c); }
  static kmeansErrorStatus initialize(H h, int alg, uint64_t seed, const double *c) {
    return kmeansInitializeD(h, alg, seed, c);
  }
  static kmeansErrorStatus assign(H h) { return kmeansAssignD(h); }
  static kmeansErrorStatus update(H h) { return kmeansUpdateCentersD(h); }
  static kmeansErrorStatus compactness(H h, double *o) { return kmeansCompactnessD(h, o); }
  static kmeansErrorStatus run(H h, int it, double c) { return kmeansRunD(h, it, c); }
  static kmeansErrorStatus miniBatch(H h, long rows) { return kmeansSetMiniBatchD(h, rows); }
  static kmeansErrorStatus centers(H h, double *c) { return kmeansGetCentersD(h, c); }
  static kmeansErrorStatus labels(H h, int *l) { return kmeansGetLabelsD(h, l); }
  static kmeansErrorStatus stats(H h, kmeansStats *s) { return kmeansGetStatsD(h, s); }
};

}  // namespace detail

template <class TYPE>
class kmeans {
  typedef detail::Api<TYPE> A;

 public:
  kmeans(const int m, const int n, const int k, const TYPE criterion, const int maxIters,
         const int numRetries, const int initAlgorithm, const int doFullSGEMM,
         const int matrixFormat, const TYPE miniBatchFraction = 1.0)
      : m_(m), n_(n), k_(k), criterion_(criterion), maxIters_(maxIters),
        initAlgorithm_(initAlgorithm), fraction_(miniBatchFraction) {
    (void)numRetries;
    (void)matrixFormat;
    kmeansOptions opt = {};
    opt.full_gemm = doFullSGEMM ? 1 : 0;
    status_ = A::create(&h_, m, n, k, &opt);
  }
  virtual ~kmeans() {
    if (h_) A::destroy(h_);
  }
  kmeans(const kmeans &) = delete;
  kmeans &operator=(const kmeans &) = delete;

  /* kmeans.hpp:198-200: converged when relErr <= criterion */
  virtual kmeansErrorStatus getConvergenceFlag() {
    kmeansStats s;
    if (!h_ || A::stats(h_, &s) != KMEANS_SUCCESS) return KMEANS_ERROR_COMPUTE;
    return s.rel_err <= criterion_ ? KMEANS_SUCCESS : KMEANS_NONCONVERGENCE;
  }
  /* result: n x k (D x K), as the reference writes it */
  virtual kmeansErrorStatus getClusterCenters(TYPE *result) {
    if (!h_ || !result) return KMEANS_ERROR_INVALID_ARGUMENT;
    std::vector<TYPE> c((size_t)k_ * n_);
    kmeansErrorStatus st = A::centers(h_, c.data());
    if (st != KMEANS_SUCCESS) return st;
    for (int j = 0; j < k_; ++j)
      for (int q = 0; q < n_; ++q) result[(size_t)q * k_ + j] = c[(size_t)j * n_ + q];
    return KMEANS_SUCCESS;
  }
  /* data: m x n host, C-ordered (copied to the GPU) */
  virtual kmeansErrorStatus buildData(const TYPE *data) {
    return h_ ? A::setData(h_, data) : status_;
  }
  /* initAlgorithm 1: result (n x k) holds the initial centres; 2: k-means++;
   * 3: k rows sampled with seed 42; otherwise the rows the reference's own
   * initialize picks (kmeansReferenceInitRows).  The initial centres are
   * written to result. */
  virtual kmeansErrorStatus initialize(TYPE *result) {
    if (!h_) return status_;
    kmeansErrorStatus st = A::initialize(h_, initAlgorithm_, 42, result);
    if (st != KMEANS_SUCCESS || !result) return st;
    return getClusterCenters(result);
  }
  virtual kmeansErrorStatus constructMiniBatch(int index) {
    (void)index;  // the engine walks the mini-batch windows itself
    return KMEANS_SUCCESS;
  }
  virtual kmeansErrorStatus compute() {
    return fraction_ < (TYPE)1 ? computeMiniBatch() : computeStandard();
  }
  virtual kmeansErrorStatus computeStandard() {
    if (!h_) return status_;
    kmeansErrorStatus st = A::miniBatch(h_, 0);
    return st == KMEANS_SUCCESS ? A::run(h_, maxIters_, criterion_) : st;
  }
  virtual kmeansErrorStatus computeMiniBatch() {
    if (!h_) return status_;
    long rows = (long)(fraction_ * m_);
    if (rows < k_) rows = k_ < m_ ? k_ : m_;
    kmeansErrorStatus st = A::miniBatch(h_, rows);
    return st == KMEANS_SUCCESS ? A::run(h_, maxIters_, criterion_) : st;
  }
  virtual kmeansErrorStatus closestCenters() { return h_ ? A::assign(h_) : status_; }
  virtual kmeansErrorStatus clusterCenters() { return h_ ? A::update(h_) : status_; }
  virtual kmeansErrorStatus compactness() {
    return h_ ? A::compactness(h_, &compactness_) : status_;
  }
  virtual kmeansErrorStatus transpose() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus reset() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus computeTiling() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus copyTileToDevice(bool standardApproach = true) {
    (void)standardApproach;
    return KMEANS_SUCCESS;
  }

  /* additions */
  kmeansErrorStatus getLabels(int *labels) { return h_ ? A::labels(h_, labels) : status_; }
  /* the last compactness(): from compactness() if called, else the loop's */
  double getCompactness() {
    if (compactness_ >= 0.0) return compactness_;
    kmeansStats s;
    return h_ && A::stats(h_, &s) == KMEANS_SUCCESS ? s.compactness : -1.0;
  }
  /* centers: k x n host */
  kmeansErrorStatus setInitialCenters(const TYPE *centers) {
    return h_ ? A::setCenters(h_, centers) : status_;
  }
  typename A::H handle() const { return h_; }

 private:
  typename A::H h_ = nullptr;
  kmeansErrorStatus status_ = KMEANS_SUCCESS;
  int m_, n_, k_;
  TYPE criterion_;
  int maxIters_, initAlgorithm_;
  TYPE fraction_;
  double compactness_ = -1.0;
};

}  // namespace kmeans_amd

#endif /* KMEANS_AMD_HPP_ */
