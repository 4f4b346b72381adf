// Library-GEMM assignment: the doFullSGEMM = 1 path of the reference
// (kmeans.cpp:918-992: cublasSgemm / cublasDgemm of the centres with the
// tile of points into an m x k product in device memory, then
// rowTransformMinimumF/D, RowTransformMinimum.cu:3-58, for the argmin of
// |x|^2 + |c|^2 - 2 x.c).
//
// Here the GEMM is rocBLAS SGEMM / DGEMM (its own summation order) and the
// argmin is this library's k_row_transform_min.  The labels therefore follow
// the library's rounding, not the canonical fma chain of the fused path:
// they agree with it except where two centres are within the GEMM's rounding
// of each other.  It is kept as the reference's second algorithm and as a
// measured comparison point for the fused kernels (the product round trip
// through HBM is what the fused kernels avoid).
#include <rocblas/rocblas.h>

#include <algorithm>

#include "km_engine.hpp"
#include "km_gemm.hpp"

namespace km {

struct GemmAssign {
  rocblas_handle h = nullptr;
  void *P = nullptr;   // tile x k product
  void *xn = nullptr;  // tile row norms (when the caller has none)
  long tile = 0;
};

void gemm_destroy(GemmAssign *g) {
  if (!g) return;
  if (g->h) rocblas_destroy_handle(g->h);
  hipFree(g->P);
  hipFree(g->xn);
  delete g;
}

// rows of one GEMM: the product buffer holds at most 2^29 elements
static long gemm_tile(long m, int k) {
  const long t = std::max(64L, ((1L << 29) / k) & ~63L);
  return std::min(m, t);
}

static hipError_t gemm_setup(GemmAssign **pg, hipStream_t s, long rows, int k, size_t esz) {
  GemmAssign *g = *pg;
  if (!g) {
    g = new GemmAssign();
    *pg = g;
    if (rocblas_create_handle(&g->h) != rocblas_status_success) return hipErrorUnknown;
    if (rocblas_set_pointer_mode(g->h, rocblas_pointer_mode_host) != rocblas_status_success)
      return hipErrorUnknown;
  }
  if (rocblas_set_stream(g->h, s) != rocblas_status_success) return hipErrorUnknown;
  const long t = gemm_tile(rows, k);
  if (t > g->tile) {
    hipFree(g->P);
    hipFree(g->xn);
    g->P = nullptr;
    g->xn = nullptr;
    g->tile = 0;
    hipError_t err;
    if ((err = hipMalloc(&g->P, esz * (size_t)t * k)) != hipSuccess) return err;
    if ((err = hipMalloc(&g->xn, esz * (size_t)t)) != hipSuccess) return err;
    g->tile = t;
  }
  return hipSuccess;
}

static rocblas_status gemm_call(rocblas_handle h, int k, long tr, int n, const float *Ct, int ldc,
                                const float *X, float *P) {
  const float one = 1.0f, zero = 0.0f;
  return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, k, (rocblas_int)tr, n,
                       &one, Ct, ldc, X, n, &zero, P, k);
}
static rocblas_status gemm_call(rocblas_handle h, int k, long tr, int n, const double *Ct, int ldc,
                                const double *X, double *P) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, k, (rocblas_int)tr, n,
                       &one, Ct, ldc, X, n, &zero, P, k);
}

template <typename T>
static hipError_t gemm_assign_t(GemmAssign **pg, hipStream_t s, const T *X, const T *xn_rows,
                                long rows, int n, int k, const T *Ct, int ldc, const T *nc,
                                int *labels) {
  hipError_t err;
  if (rows <= 0) return hipSuccess;
  if ((err = gemm_setup(pg, s, rows, k, sizeof(T))) != hipSuccess) return err;
  GemmAssign *g = *pg;
  T *P = (T *)g->P;
  for (long r = 0; r < rows; r += g->tile) {
    const long tr = std::min(g->tile, rows - r);
    const T *Xr = X + r * n;
    // P (k x tr, column-major) = Ct^T (k x n; Ct is n x ldc row-major) * X^T
    if (gemm_call(g->h, k, tr, n, Ct, ldc, Xr, P) != rocblas_status_success)
      return hipErrorLaunchFailure;
    const T *xn = xn_rows ? xn_rows + r : (const T *)g->xn;
    if (!xn_rows && (err = launch_row_norms(s, tr, n, Xr, (T *)g->xn)) != hipSuccess) return err;
    // rows with no value below FLT_MAX (NaN data) get label 0, a valid centre
    if ((err = launch_row_transform_min(s, 0, tr, k, xn, nc, P, labels + r, 0)) != hipSuccess)
      return err;
  }
  return hipSuccess;
}

hipError_t gemm_assign(GemmAssign **g, hipStream_t s, const float *X, const float *xn, long rows,
                       int n, int k, const float *Ct, int ldc, const float *nc, int *labels) {
  return gemm_assign_t(g, s, X, xn, rows, n, k, Ct, ldc, nc, labels);
}
hipError_t gemm_assign(GemmAssign **g, hipStream_t s, const double *X, const double *xn,
                       long rows, int n, int k, const double *Ct, int ldc, const double *nc,
                       int *labels) {
  return gemm_assign_t(g, s, X, xn, rows, n, k, Ct, ldc, nc, labels);
}

hipError_t gemm_assign(Engine *e, const float *X, const float *xn, long rows, int *labels) {
  return gemm_assign_t((GemmAssign **)&e->gemm, e->stream, X, xn, rows, e->n, e->k, e->dCt,
                       e->ldc, e->dnc, labels);
}

void gemm_free(Engine *e) {
  gemm_destroy((GemmAssign *)e->gemm);
  e->gemm = nullptr;
}

}  // namespace km
