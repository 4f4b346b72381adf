// MLP proposer kernels (BF16 MFMA products, FP32 accumulation and masters).
#pragma once
#include "km_common.hpp"

namespace km {
// FP32 -> BF16 bits, round to nearest even (NaN stays NaN)
__device__ inline uint16_t mlp_bf16(float f) { return bf16_rne(f); }

struct MlpShape {
  int din, h, g;  // input width (multiple of 16), hidden width, outputs
};
size_t mlp_param_count(const MlpShape &sh);
void mlp_offsets(const MlpShape &sh, size_t off[6]);
size_t mlp_ws_bytes(const MlpShape &sh, int B);
hipError_t gemm_nt(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda,
                   const uint16_t *Bt, int ldb, const float *bias, bool relu, float *C, int ldc,
                   uint16_t *Ch, int ldh);
hipError_t transpose_b16(hipStream_t s, int rows, int cols, const uint16_t *in, int ld_in,
                         uint16_t *out, int ld_out);
hipError_t mlp_fwd_bwd(hipStream_t s, const MlpShape &sh, int B, const uint16_t *X, const int *t,
                       const float *P, float *G, float *loss, void *ws);
hipError_t mlp_adam(hipStream_t s, size_t n, float *P, const float *G, float *m, float *v, float lr,
                    int step, uint16_t *Ph);
hipError_t mlp_forward(hipStream_t s, const MlpShape &sh, int rows, const uint16_t *X, int ldx,
                       const uint16_t *Ph, const float *P, uint16_t *H1, uint16_t *H2, float *Z);
hipError_t params_to_bf16(hipStream_t s, size_t n, const float *P, uint16_t *Ph);
}  // namespace km
