// Library-GEMM assignment (rocBLAS SGEMM / DGEMM + row argmin), the
// reference's doFullSGEMM = 1 path.
#pragma once
#include "km_common.hpp"

namespace km {
struct Engine;
struct GemmAssign;
// labels[0..rows) of the rows X (rows x n) against centres Ct (n x ldc) with
// norms nc; xn (nullable): the rows' squared norms, computed tile by tile when
// null; *g holds the library handle and product buffer (created on first use,
// freed by gemm_destroy)
hipError_t gemm_assign(GemmAssign **g, hipStream_t s, const float *X, const float *xn, long rows,
                       int n, int k, const float *Ct, int ldc, const float *nc, int *labels);
hipError_t gemm_assign(GemmAssign **g, hipStream_t s, const double *X, const double *xn,
                       long rows, int n, int k, const double *Ct, int ldc, const double *nc,
                       int *labels);
void gemm_destroy(GemmAssign *g);
// the FP32 engine's form (its centres, its stream)
hipError_t gemm_assign(Engine *e, const float *X, const float *xn, long rows, int *labels);
void gemm_free(Engine *e);
}  // namespace km
