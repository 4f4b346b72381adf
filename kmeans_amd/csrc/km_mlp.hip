// Small MLP (candidate-group proposer) on BF16 MFMA: forward, backward and
// Adam, written for this engine (no library GEMM).  Layers are
// y = x W^T + b with W stored out x in (the same orientation as the weight
// of a linear layer), ReLU between layers, softmax cross-entropy on the
// logits.  Operands of every product are BF16, accumulation is FP32, master
// weights and Adam moments are FP32.
#include <float.h>
#include <math.h>

#include "km_common.hpp"
#include "km_mlp.hpp"

namespace km {

typedef __attribute__((ext_vector_type(8))) __bf16 mlp_bf16x8;
typedef __attribute__((ext_vector_type(16))) float mlp_f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int mlp_u32x4;

__device__ inline float mlp_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline int mlp_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// C[M x N] = A[M x K] * Bt[N x K]^T (+ bias[N]) (ReLU); A and Bt BF16 with
// K contiguous (K a multiple of 16, rows padded with zeros by the caller).
// Outputs: FP32 C (ldc) and/or BF16 Ch (ldh).  Each wave computes a 64 x 64
// block (2 x 2 MFMA tiles sharing fragments), 4 waves per 128 x 128 block.
__global__ __launch_bounds__(256) void k_gemm_nt(int M, int N, int K, const uint16_t *__restrict__ A,
                                                 int lda, const uint16_t *__restrict__ Bt, int ldb,
                                                 const float *__restrict__ bias, int relu,
                                                 float *__restrict__ C, int ldc,
                                                 uint16_t *__restrict__ Ch, int ldh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
  const int m0 = blockIdx.y * 128 + (w >> 1) * 64, n0 = blockIdx.x * 128 + (w & 1) * 64;
  mlp_f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  const int ar0 = m0 + col, ar1 = m0 + 32 + col, bc0 = n0 + col, bc1 = n0 + 32 + col;
  const mlp_u32x4 z = {0u, 0u, 0u, 0u};
  for (int k = 0; k < K; k += 16) {
    const int kk = k + 8 * h;
    const mlp_u32x4 a0 = ar0 < M ? *reinterpret_cast<const mlp_u32x4 *>(A + (long)ar0 * lda + kk) : z;
    const mlp_u32x4 a1 = ar1 < M ? *reinterpret_cast<const mlp_u32x4 *>(A + (long)ar1 * lda + kk) : z;
    const mlp_u32x4 b0 = bc0 < N ? *reinterpret_cast<const mlp_u32x4 *>(Bt + (long)bc0 * ldb + kk) : z;
    const mlp_u32x4 b1 = bc1 < N ? *reinterpret_cast<const mlp_u32x4 *>(Bt + (long)bc1 * ldb + kk) : z;
    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mlp_bf16x8, a0),
                                                        __builtin_bit_cast(mlp_bf16x8, b0), acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mlp_bf16x8, a0),
                                                        __builtin_bit_cast(mlp_bf16x8, b1), acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mlp_bf16x8, a1),
                                                        __builtin_bit_cast(mlp_bf16x8, b0), acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mlp_bf16x8, a1),
                                                        __builtin_bit_cast(mlp_bf16x8, b1), acc[1][1], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int cc = n0 + 32 * b + col;
      if (cc >= N) continue;
      const float bv = bias ? bias[cc] : 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rr = m0 + 32 * a + mlp_row(r, h);
        if (rr >= M) continue;
        float v = acc[a][b][r] + bv;
        if (relu) v = fmaxf(v, 0.0f);
        if (C) C[(long)rr * ldc + cc] = v;
        if (Ch) Ch[(long)rr * ldh + cc] = mlp_bf16(v);
      }
    }
}

// out[c][r] = in[r][c] (BF16), out zero-padded to ld_out columns.
__global__ void k_transpose_b16(int rows, int cols, const uint16_t *__restrict__ in, int ld_in,
                                uint16_t *__restrict__ out, int ld_out) {
  __shared__ uint16_t t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    t[i][tx] = (r < rows && c < cols) ? in[(long)r * ld_in + c] : (uint16_t)0;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (c < cols && r < ld_out) out[(long)c * ld_out + r] = t[tx][i];
  }
}

__global__ void k_to_bf16(long n, const float *__restrict__ in, uint16_t *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mlp_bf16(in[i]);
}

// Softmax cross-entropy, one wave per row: dZ = (softmax(Z) - onehot(t)) / B
// (FP32 and BF16 copies), per-row loss into loss[row].
__global__ __launch_bounds__(64) void k_softmax_ce(int B, int G, const float *__restrict__ Z,
                                                   const int *__restrict__ t,
                                                   float *__restrict__ dZ,
                                                   uint16_t *__restrict__ dZh,
                                                   float *__restrict__ loss) {
  const int row = blockIdx.x, lane = threadIdx.x;
  if (row >= B) return;
  const float *z = Z + (long)row * G;
  float mx = -INFINITY;
  for (int g = lane; g < G; g += 64) mx = fmaxf(mx, z[g]);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  float se = 0.0f;
  for (int g = lane; g < G; g += 64) se += expf(z[g] - mx);
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
  const int tg = t[row];
  const bool tok = tg >= 0 && tg < G;
  const float inv = 1.0f / (float)B;
  for (int g = lane; g < G; g += 64) {
    const float p = expf(z[g] - mx) / se;
    const float d = (p - (g == tg ? 1.0f : 0.0f)) * inv;
    dZ[(long)row * G + g] = d;
    dZh[(long)row * G + g] = mlp_bf16(d);
  }
  if (lane == 0) loss[row] = tok ? logf(se) + mx - z[tg] : __builtin_nanf("");
}

// dH = dY where H > 0 (ReLU backward); FP32 and BF16 copies.
__global__ void k_relu_bwd(long n, const float *__restrict__ dY, const uint16_t *__restrict__ H,
                           float *__restrict__ dH, uint16_t *__restrict__ dHh) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = mlp_f32(H[i]) > 0.0f ? dY[i] : 0.0f;
  dH[i] = v;
  dHh[i] = mlp_bf16(v);
}

// db[c] = sum over rows of dY[r][c]: 64 columns per block, 16 row phases
// (rows r = phase mod 16, ascending) combined in phase order, so the result
// does not depend on scheduling.
__global__ __launch_bounds__(1024) void k_colsum(int rows, int cols, const float *__restrict__ dY,
                                                 float *__restrict__ db) {
  __shared__ float part[16][64];
  const int lane = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (c < cols)
    for (int r = ph; r < rows; r += 16) s += dY[(long)r * cols + c];
  part[ph][lane] = s;
  __syncthreads();
  if (ph == 0 && c < cols) {
    float t = 0.0f;
    for (int q = 0; q < 16; ++q) t += part[q][lane];
    db[c] = t;
  }
}

// Adam (Kingma & Ba, ICLR 2015) on FP32 masters, refreshes the BF16 copy.
__global__ void k_adam(long n, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t *__restrict__ wh) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wh) wh[i] = mlp_bf16(wi);
}

__global__ void k_scale(float *x, float a) { *x *= a; }

// ------------------------------------------------------------------ host
hipError_t gemm_nt(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda,
                   const uint16_t *Bt, int ldb, const float *bias, bool relu, float *C, int ldc,
                   uint16_t *Ch, int ldh) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 g((N + 127) / 128, (M + 127) / 128);
  hipLaunchKernelGGL(k_gemm_nt, g, dim3(256), 0, s, M, N, K, A, lda, Bt, ldb, bias, relu ? 1 : 0,
                     C, ldc, Ch, ldh);
  return hipGetLastError();
}

hipError_t transpose_b16(hipStream_t s, int rows, int cols, const uint16_t *in, int ld_in,
                         uint16_t *out, int ld_out) {
  dim3 g((cols + 31) / 32, (ld_out + 31) / 32);
  hipLaunchKernelGGL(k_transpose_b16, g, dim3(256), 0, s, rows, cols, in, ld_in, out, ld_out);
  return hipGetLastError();
}

static inline unsigned grid1(long n) { return (unsigned)((n + 255) / 256); }

size_t mlp_param_count(const MlpShape &sh) {
  return (size_t)sh.h * sh.din + sh.h + (size_t)sh.h * sh.h + sh.h + (size_t)sh.g * sh.h + sh.g;
}

// Parameter blocks in the flat vector: W1 (h x din), b1, W2 (h x h), b2, W3 (g x h), b3.
void mlp_offsets(const MlpShape &sh, size_t off[6]) {
  off[0] = 0;
  off[1] = off[0] + (size_t)sh.h * sh.din;
  off[2] = off[1] + sh.h;
  off[3] = off[2] + (size_t)sh.h * sh.h;
  off[4] = off[3] + sh.h;
  off[5] = off[4] + (size_t)sh.g * sh.h;
}

size_t mlp_ws_bytes(const MlpShape &sh, int B) {
  const size_t P = mlp_param_count(sh);
  const size_t Bp = ((size_t)B + 15) / 16 * 16;
  size_t b = 0;
  b += P * 2;                                  // BF16 copy of the parameters
  b += (size_t)B * sh.h * 2 * 2;               // H1, H2 (BF16)
  b += (size_t)B * sh.g * 4 * 2;               // Z, dZ
  b += (size_t)B * sh.g * 2;                   // dZ BF16
  b += (size_t)B * sh.h * 4 * 2;               // dH2 / dH1 FP32 (before and after mask)
  b += (size_t)B * sh.h * 2;                   // masked BF16
  b += (size_t)Bp * (sh.g + 2 * sh.h + sh.din) * 2;  // transposes (K = batch, padded)
  b += (size_t)sh.h * sh.h * 2 * 2 + (size_t)sh.g * sh.h * 2;  // W transposes
  b += (size_t)B * 4;                          // per-row loss
  return b + 64 * 256;
}

// One forward + backward pass on a batch: X (B x din BF16, din multiple of
// 16), targets t (B), parameters P (FP32 flat), gradients G (FP32 flat),
// loss (device float, mean over the batch).  Workspace from mlp_ws_bytes.
hipError_t mlp_fwd_bwd(hipStream_t s, const MlpShape &sh, int B, const uint16_t *X, const int *t,
                       const float *P, float *G, float *loss, void *ws) {
  hipError_t e;
  size_t off[6];
  mlp_offsets(sh, off);
  const size_t Pn = mlp_param_count(sh);
  const int Bp = (B + 15) / 16 * 16;
  char *p = (char *)ws;
  auto take = [&](size_t bytes) {
    void *r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
  };
  uint16_t *Ph = (uint16_t *)take(Pn * 2);
  uint16_t *H1 = (uint16_t *)take((size_t)B * sh.h * 2);
  uint16_t *H2 = (uint16_t *)take((size_t)B * sh.h * 2);
  float *Z = (float *)take((size_t)B * sh.g * 4);
  float *dZ = (float *)take((size_t)B * sh.g * 4);
  uint16_t *dZh = (uint16_t *)take((size_t)B * sh.g * 2);
  float *dH = (float *)take((size_t)B * sh.h * 4);
  float *dHm = (float *)take((size_t)B * sh.h * 4);
  uint16_t *dHh = (uint16_t *)take((size_t)B * sh.h * 2);
  uint16_t *Tz = (uint16_t *)take((size_t)Bp * sh.g * 2);    // dZ^T  (g x Bp)
  uint16_t *Ta = (uint16_t *)take((size_t)Bp * sh.h * 2);    // H2^T / H1^T (h x Bp)
  uint16_t *Tb = (uint16_t *)take((size_t)Bp * sh.h * 2);    // dH^T (h x Bp)
  uint16_t *Tx = (uint16_t *)take((size_t)Bp * sh.din * 2);  // X^T (din x Bp)
  uint16_t *W3t = (uint16_t *)take((size_t)sh.g * sh.h * 2);  // h x g
  uint16_t *W2t = (uint16_t *)take((size_t)sh.h * sh.h * 2);  // h x h
  float *rl = (float *)take((size_t)B * 4);
  const uint16_t *W1h = Ph + off[0], *W2h = Ph + off[2], *W3h = Ph + off[4];
  const float *b1 = P + off[1], *b2 = P + off[3], *b3 = P + off[5];
  float *gW1 = G + off[0], *gb1 = G + off[1], *gW2 = G + off[2], *gb2 = G + off[3];
  float *gW3 = G + off[4], *gb3 = G + off[5];

  hipLaunchKernelGGL(k_to_bf16, dim3(grid1(Pn)), dim3(256), 0, s, (long)Pn, P, Ph);
  // forward
  if ((e = gemm_nt(s, B, sh.h, sh.din, X, sh.din, W1h, sh.din, b1, true, nullptr, 0, H1, sh.h))) return e;
  if ((e = gemm_nt(s, B, sh.h, sh.h, H1, sh.h, W2h, sh.h, b2, true, nullptr, 0, H2, sh.h))) return e;
  if ((e = gemm_nt(s, B, sh.g, sh.h, H2, sh.h, W3h, sh.h, b3, false, Z, sh.g, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_softmax_ce, dim3(B), dim3(64), 0, s, B, sh.g, Z, t, dZ, dZh, rl);
  // layer 3: dW3 = dZ^T H2, db3 = colsum dZ, dH2 = dZ W3
  if ((e = transpose_b16(s, B, sh.g, dZh, sh.g, Tz, Bp))) return e;
  if ((e = transpose_b16(s, B, sh.h, H2, sh.h, Ta, Bp))) return e;
  if ((e = gemm_nt(s, sh.g, sh.h, Bp, Tz, Bp, Ta, Bp, nullptr, false, gW3, sh.h, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_colsum, dim3((sh.g + 63) / 64), dim3(1024), 0, s, B, sh.g, dZ, gb3);
  if ((e = transpose_b16(s, sh.g, sh.h, W3h, sh.h, W3t, sh.g))) return e;
  if ((e = gemm_nt(s, B, sh.h, sh.g, dZh, sh.g, W3t, sh.g, nullptr, false, dH, sh.h, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1((long)B * sh.h)), dim3(256), 0, s, (long)B * sh.h, dH,
                     H2, dHm, dHh);
  // layer 2: dW2 = dH2^T H1, db2, dH1 = dH2 W2
  if ((e = transpose_b16(s, B, sh.h, dHh, sh.h, Tb, Bp))) return e;
  if ((e = transpose_b16(s, B, sh.h, H1, sh.h, Ta, Bp))) return e;
  if ((e = gemm_nt(s, sh.h, sh.h, Bp, Tb, Bp, Ta, Bp, nullptr, false, gW2, sh.h, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_colsum, dim3((sh.h + 63) / 64), dim3(1024), 0, s, B, sh.h, dHm, gb2);
  if ((e = transpose_b16(s, sh.h, sh.h, W2h, sh.h, W2t, sh.h))) return e;
  if ((e = gemm_nt(s, B, sh.h, sh.h, dHh, sh.h, W2t, sh.h, nullptr, false, dH, sh.h, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1((long)B * sh.h)), dim3(256), 0, s, (long)B * sh.h, dH,
                     H1, dHm, dHh);
  // layer 1: dW1 = dH1^T X, db1
  if ((e = transpose_b16(s, B, sh.h, dHh, sh.h, Tb, Bp))) return e;
  if ((e = transpose_b16(s, B, sh.din, X, sh.din, Tx, Bp))) return e;
  if ((e = gemm_nt(s, sh.h, sh.din, Bp, Tb, Bp, Tx, Bp, nullptr, false, gW1, sh.din, nullptr, 0))) return e;
  hipLaunchKernelGGL(k_colsum, dim3((sh.h + 63) / 64), dim3(1024), 0, s, B, sh.h, dHm, gb1);
  // mean loss (left to right over the batch rows)
  if (loss) {
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(1024), 0, s, B, 1, rl, loss);
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(1), 0, s, loss, 1.0f / (float)B);
  }
  return hipGetLastError();
}

hipError_t mlp_adam(hipStream_t s, size_t n, float *P, const float *G, float *m, float *v, float lr,
                    int step, uint16_t *Ph) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1((long)n)), dim3(256), 0, s, (long)n, P, G, m, v, lr, b1, b2,
                     eps, c1, c2, Ph);
  return hipGetLastError();
}

// Inference: logits (FP32, rows x g) for BF16 inputs, chunked by the caller.
hipError_t mlp_forward(hipStream_t s, const MlpShape &sh, int rows, const uint16_t *X, int ldx,
                       const uint16_t *Ph, const float *P, uint16_t *H1, uint16_t *H2, float *Z) {
  hipError_t e;
  size_t off[6];
  mlp_offsets(sh, off);
  if ((e = gemm_nt(s, rows, sh.h, sh.din, X, ldx, Ph + off[0], sh.din, P + off[1], true, nullptr, 0,
                   H1, sh.h)))
    return e;
  if ((e = gemm_nt(s, rows, sh.h, sh.h, H1, sh.h, Ph + off[2], sh.h, P + off[3], true, nullptr, 0, H2,
                   sh.h)))
    return e;
  return gemm_nt(s, rows, sh.g, sh.h, H2, sh.h, Ph + off[4], sh.h, P + off[5], false, Z, sh.g,
                 nullptr, 0);
}

hipError_t params_to_bf16(hipStream_t s, size_t n, const float *P, uint16_t *Ph) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1((long)n)), dim3(256), 0, s, (long)n, P, Ph);
  return hipGetLastError();
}

}  // namespace km
