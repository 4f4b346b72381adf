/*
 * kmeans_amd.hpp -- the reference's host class kmeans<TYPE>
 * (main/cpp/kmeans.hpp:174-312) over the C-ABI of kmeans_amd.h, header only.
 *
 * Same constructor arguments and public method names, so code written
 * against the reference's class compiles against this one after the include
 * and namespace change.  kmeans<float> drives the engine handle step by step
 * (closestCenters / clusterCenters / compactness are one C call each);
 * kmeans<double> runs its loop through computeKmeansD, so its step methods
 * return KMEANS_ERROR_COMPUTE.  Methods whose work the engine does
 * implicitly (transpose, reset, computeTiling, copyTileToDevice,
 * constructMiniBatch) return KMEANS_SUCCESS.  Added beyond the reference:
 * getLabels, getCompactness, setInitialCenters.
 */
#ifndef KMEANS_AMD_HPP_
#define KMEANS_AMD_HPP_

#include <stdint.h>

#include <vector>

#include "kmeans_amd.h"

namespace kmeans_amd {

template <class TYPE>
class kmeans;

template <>
class kmeans<float> {
 public:
  kmeans(const int m, const int n, const int k, const float criterion, const int maxIters,
         const int numRetries, const int initAlgorithm, const int doFullSGEMM,
         const int matrixFormat, const float miniBatchFraction = 1.0f)
      : m_(m), n_(n), k_(k), criterion_(criterion), maxIters_(maxIters),
        initAlgorithm_(initAlgorithm), fraction_(miniBatchFraction) {
    (void)numRetries;
    (void)matrixFormat;
    kmeansOptions opt = {};
    opt.full_gemm = doFullSGEMM ? 1 : 0;
    status_ = kmeansCreate(&h_, m, n, k, &opt);
  }
  virtual ~kmeans() {
    if (h_) kmeansDestroy(h_);
  }
  kmeans(const kmeans &) = delete;
  kmeans &operator=(const kmeans &) = delete;

  virtual kmeansErrorStatus getConvergenceFlag() {
    kmeansStats s;
    if (!h_ || kmeansGetStats(h_, &s) != KMEANS_SUCCESS) return KMEANS_ERROR_COMPUTE;
    return s.rel_err <= criterion_ ? KMEANS_SUCCESS : KMEANS_NONCONVERGENCE;
  }
  /* result: n x k (D x K), as the reference writes it */
  virtual kmeansErrorStatus getClusterCenters(float *result) {
    if (!h_ || !result) return KMEANS_ERROR_INVALID_ARGUMENT;
    std::vector<float> c((size_t)k_ * n_);
    kmeansErrorStatus st = kmeansGetCenters(h_, c.data());
    if (st != KMEANS_SUCCESS) return st;
    for (int j = 0; j < k_; ++j)
      for (int q = 0; q < n_; ++q) result[(size_t)q * k_ + j] = c[(size_t)j * n_ + q];
    return KMEANS_SUCCESS;
  }
  /* data: m x n host, C-ordered (copied to the GPU) */
  virtual kmeansErrorStatus buildData(const float *data) {
    return h_ ? kmeansSetData(h_, data, 0) : status_;
  }
  /* initAlgorithm 1: result (n x k) holds the initial centres; 2: k-means++;
   * otherwise k sampled rows.  The initial centres are written to result. */
  virtual kmeansErrorStatus initialize(float *result) {
    if (!h_) return status_;
    kmeansErrorStatus st = kmeansInitialize(h_, initAlgorithm_, 42, result);
    if (st != KMEANS_SUCCESS || !result) return st;
    return getClusterCenters(result);
  }
  virtual kmeansErrorStatus constructMiniBatch(int index) {
    (void)index;  // the engine walks the mini-batch windows itself
    return KMEANS_SUCCESS;
  }
  virtual kmeansErrorStatus compute() {
    return fraction_ < 1.0f ? computeMiniBatch() : computeStandard();
  }
  virtual kmeansErrorStatus computeStandard() {
    if (!h_) return status_;
    kmeansErrorStatus st = kmeansSetMiniBatch(h_, 0);
    return st == KMEANS_SUCCESS ? kmeansRun(h_, maxIters_, criterion_) : st;
  }
  virtual kmeansErrorStatus computeMiniBatch() {
    if (!h_) return status_;
    long rows = (long)(fraction_ * m_);
    if (rows < k_) rows = k_ < m_ ? k_ : m_;
    kmeansErrorStatus st = kmeansSetMiniBatch(h_, rows);
    return st == KMEANS_SUCCESS ? kmeansRun(h_, maxIters_, criterion_) : st;
  }
  virtual kmeansErrorStatus closestCenters() { return h_ ? kmeansAssign(h_) : status_; }
  virtual kmeansErrorStatus clusterCenters() { return h_ ? kmeansUpdateCenters(h_) : status_; }
  virtual kmeansErrorStatus compactness() {
    return h_ ? kmeansCompactness(h_, &compactness_) : status_;
  }
  virtual kmeansErrorStatus transpose() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus reset() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus computeTiling() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus copyTileToDevice(bool standardApproach = true) {
    (void)standardApproach;
    return KMEANS_SUCCESS;
  }

  /* additions */
  kmeansErrorStatus getLabels(int *labels) { return h_ ? kmeansGetLabels(h_, labels) : status_; }
  /* the last compactness(): from compactness() if called, else the loop's */
  double getCompactness() {
    if (compactness_ >= 0.0) return compactness_;
    kmeansStats s;
    return h_ && kmeansGetStats(h_, &s) == KMEANS_SUCCESS ? s.compactness : -1.0;
  }
  /* centers: k x n host */
  kmeansErrorStatus setInitialCenters(const float *centers) {
    return h_ ? kmeansSetCenters(h_, centers) : status_;
  }
  kmeansHandle handle() const { return h_; }

 private:
  kmeansHandle h_ = nullptr;
  kmeansErrorStatus status_ = KMEANS_SUCCESS;
  int m_, n_, k_;
  float criterion_;
  int maxIters_, initAlgorithm_;
  float fraction_;
  double compactness_ = -1.0;
};

template <>
class kmeans<double> {
 public:
  kmeans(const int m, const int n, const int k, const double criterion, const int maxIters,
         const int numRetries, const int initAlgorithm, const int doFullSGEMM,
         const int matrixFormat, const double miniBatchFraction = 1.0)
      : m_(m), n_(n), k_(k), criterion_(criterion), maxIters_(maxIters),
        numRetries_(numRetries), initAlgorithm_(initAlgorithm), doFullSGEMM_(doFullSGEMM),
        matrixFormat_(matrixFormat), fraction_(miniBatchFraction),
        centers_((size_t)k * n, 0.0) {}
  virtual ~kmeans() {}

  virtual kmeansErrorStatus getConvergenceFlag() { return status_; }
  virtual kmeansErrorStatus getClusterCenters(double *result) {
    if (!result) return KMEANS_ERROR_INVALID_ARGUMENT;
    for (size_t i = 0; i < centers_.size(); ++i) result[i] = centers_[i];
    return KMEANS_SUCCESS;
  }
  /* the data must stay valid until compute() */
  virtual kmeansErrorStatus buildData(const double *data) {
    data_ = data;
    return data ? KMEANS_SUCCESS : KMEANS_ERROR_INVALID_ARGUMENT;
  }
  /* initAlgorithm 1: result (n x k) holds the initial centres, kept for
   * compute(); other values are applied inside computeKmeansD. */
  virtual kmeansErrorStatus initialize(double *result) {
    if (initAlgorithm_ == 1 && result)
      for (size_t i = 0; i < centers_.size(); ++i) centers_[i] = result[i];
    return KMEANS_SUCCESS;
  }
  virtual kmeansErrorStatus constructMiniBatch(int index) {
    (void)index;
    return KMEANS_SUCCESS;
  }
  virtual kmeansErrorStatus compute() {
    if (!data_) return KMEANS_ERROR_BUILDDATA;
    status_ = computeKmeansD(data_, m_, n_, k_, criterion_, maxIters_, numRetries_,
                             initAlgorithm_, doFullSGEMM_, matrixFormat_, centers_.data(),
                             fraction_);
    return status_ == KMEANS_NONCONVERGENCE ? KMEANS_SUCCESS : status_;
  }
  virtual kmeansErrorStatus computeStandard() { return compute(); }
  virtual kmeansErrorStatus computeMiniBatch() { return compute(); }
  /* the double loop runs inside computeKmeansD: no separate steps */
  virtual kmeansErrorStatus closestCenters() { return KMEANS_ERROR_COMPUTE; }
  virtual kmeansErrorStatus clusterCenters() { return KMEANS_ERROR_COMPUTE; }
  virtual kmeansErrorStatus compactness() { return KMEANS_ERROR_COMPUTE; }
  virtual kmeansErrorStatus transpose() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus reset() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus computeTiling() { return KMEANS_SUCCESS; }
  virtual kmeansErrorStatus copyTileToDevice(bool standardApproach = true) {
    (void)standardApproach;
    return KMEANS_SUCCESS;
  }

 private:
  int m_, n_, k_;
  double criterion_;
  int maxIters_, numRetries_, initAlgorithm_, doFullSGEMM_, matrixFormat_;
  double fraction_;
  std::vector<double> centers_;
  const double *data_ = nullptr;
  kmeansErrorStatus status_ = KMEANS_NONCONVERGENCE;
};

}  // namespace kmeans_amd

#endif /* KMEANS_AMD_HPP_ */
