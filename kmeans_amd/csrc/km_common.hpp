// Shared device/host helpers for the gfx950 K-means kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/kmeans_amd.h"

#define KM_SB 16384         // points per histogram/scatter sub-block

#define KM_HIP_CHECK(expr, code)                                                     \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "kmeans_amd: %s failed at %s:%d: %s\n", #expr, __FILE__,       \
              __LINE__, hipGetErrorString(e_));                                      \
      return code;                                                                   \
    }                                                                                \
  } while (0)

#define KM_LAUNCH_CHECK(code) KM_HIP_CHECK(hipGetLastError(), code)

static inline long km_cdiv(long a, long b) { return (a + b - 1) / b; }
// Blocks for a grid-stride loop over `work` items: the dispatch packet holds
// the grid size in work-items as 32 bits, so element-wise launches over
// m x d items (which can exceed 2^32) are capped and loop instead.
static inline unsigned km_grid(long work, int block) {
  const long b = km_cdiv(work, block);
  return (unsigned)(b < (1L << 22) ? b : (1L << 22));
}

namespace km {

// FP32 -> BF16 bits, round to nearest even, by the gfx950 conversion
// instruction (v_cvt_pk_bf16_f32; NaN stays NaN, infinities and subnormals
// convert exactly as the RNE rule gives).
__device__ inline uint16_t bf16_rne(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// two values packed into one word (first in the low half), one instruction
__device__ inline uint32_t bf16x2_rne(float a, float b) {
  typedef __attribute__((ext_vector_type(2))) float f2_t;
  typedef __attribute__((ext_vector_type(2))) __bf16 b2_t;
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2_t){a, b}, b2_t));
}

// Counter-based generator shared with the oracle: SplitMix64 finaliser
// (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) applied to (seed, stream, index).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t k = mix64(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL +
                     0x632BE59BD9B4E019ULL);
  return mix64(k + i * 0x9E3779B97F4A7C15ULL);
}
__host__ __device__ inline double rng_unif(uint64_t seed, uint64_t stream, uint64_t i) {
  return ((double)(rng_u64(seed, stream, i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// k distinct row indices in [0, m), in draw order, by Floyd's algorithm on
// the counter-based generator (seed, stream 4).  Stands in for the
// reference's srand(42) / rand() % m pick (kmeans.cpp:876-888).
inline void sample_rows(uint64_t seed, long m, int k, int *idx) {
  long t = 0;
  for (long j = m - k; j < m; ++j) {
    long r = (long)(rng_u64(seed, 4, (uint64_t)t) % (uint64_t)(j + 1));
    bool dup = false;
    for (long q = 0; q < t; ++q)
      if (idx[q] == r) { dup = true; break; }
    idx[t++] = dup ? (int)j : (int)r;
  }
}

// The rows the reference's initialize picks as initial centres
// (kmeans.cpp:553-556 shuffle, :876-888 pick), as indices into the caller's
// unshuffled rows (km_engine.hip).
void reference_init_rows(long m, int k, int *rows);

// k-means++ seeding (km_init.hip): the k picked rows of X into idx (device)
hipError_t kmeanspp(hipStream_t s, long m, int d, int k, const float *X, uint64_t seed, int *idx);
hipError_t kmeanspp(hipStream_t s, long m, int d, int k, const double *X, uint64_t seed,
                    int *idx);

kmeansErrorStatus compute_d(const double *data, int m, int n, int k, double criterion,
                            int maxIters, int initAlgorithm, int fullGemm, double *result,
                            double miniBatchFraction);

// Internal launchers (implemented in km_kernels.hip); all stream-ordered.
struct UpdateWs {
  int *hist, *offs, *tsum, *cstart, *ccount, *cq, *perm;
  int *dummy;  // SCS_THREADS ints that the staged scatter's stores past a sub-block's end write
  double *P, *S, *compj;  // S: k x (d+1), column d = the squared-norm moment
  long long *cnt;
  size_t bytes;
  // nullable: an n-sized array that receives each block's stable sort by
  // label at the block's own rows (instead of perm), kept for the next assign
  int *perm_out;
  // Reuse of the previous update's chunk partials (fd null: none).  The
  // previous update of the same rows left its sort in perm_prev (n-sized, at
  // the block's rows), its per-centre ranges in cs_prev ([cstart K | ccount
  // K | cq K+1]) and its partials in P_prev; fd (K ints) receives each
  // centre's first member position at which the two sorts differ, and every
  // chunk that lies wholly before it has the same members, so the same
  // partials, bit for bit.  fd without perm_prev: everything is computed.
  const int *perm_prev = nullptr, *cs_prev = nullptr;
  const double *P_prev = nullptr;
  int *fd = nullptr;
};
size_t update_ws_layout(long n, int d, int k, char *base, UpdateWs *ws);
// Streams, events and the two per-block workspaces of the overlapped update
// (launch_update_accumulate_2s); S, cnt and compj stay with the caller.
struct UpdateOverlap {
  hipStream_t side[2];
  hipEvent_t start, done[2], fin[2], sorted[2];
  const UpdateWs *ws[2];
};
// blk (nullable): one workspace view per canonical block (the side
// workspace's sort scratch with the block's own persistent arrays)
hipError_t launch_update_accumulate_2s(hipStream_t s, const UpdateOverlap &ov, long n, int d,
                                       int k, const float *X, const int *labels, double *S,
                                       long long *cnt, const UpdateWs *blk = nullptr);
int sub_block(int k);
// out[0..L] = exclusive scan of in[0..L) (out[L] the total); tsum holds
// scan_tiles(L) + 1 ints
long scan_tiles(long L);
hipError_t launch_scan_excl(hipStream_t s, const int *in, long L, int *tsum, int *out);

// Launchers with a float and a double form (element type of the points and
// centres; the sums are FP64 in both).
#define KM_TYPED_LAUNCHERS(T)                                                                    \
  hipError_t launch_col_norms(hipStream_t s, int k, int d, const T *C, T *nc);                   \
  hipError_t launch_transpose(hipStream_t s, int k, int d, const T *C, T *Ct, int ldc);          \
  hipError_t launch_update(hipStream_t s, long n, int d, int k, const T *X, const int *labels,   \
                           const T *C_old, T *C_new, T *Ct_new, int ldc, T *nc_new,              \
                           long long *counts, int *empty, const UpdateWs &ws);                   \
  hipError_t launch_compactness(hipStream_t s, long n, int d, int k, const T *X,                 \
                                const int *labels, const T *C, double *out, const UpdateWs &ws); \
  hipError_t launch_point_sq(hipStream_t s, long m, int d, const T *X, double *q);               \
  hipError_t launch_update_accumulate(hipStream_t s, long n, int d, int k, const T *X,           \
                                      const int *labels, double *S, long long *cnt,              \
                                      const UpdateWs &ws);                                       \
  hipError_t launch_update_finalize(hipStream_t s, int d, int k, const double *S,                \
                                    const long long *cnt, const T *C_old, T *C_new, T *Ct_new,   \
                                    int ldc, T *nc_new, long long *counts, int *empty,           \
                                    double *compj);                                              \
  hipError_t launch_col_norms_dk(hipStream_t s, int k, int d, int ld, const T *C, T *nc);        \
  hipError_t launch_row_norms(hipStream_t s, long m, int d, const T *X, T *out);                 \
  hipError_t launch_row_transform_min(hipStream_t s, long m0, long m, int n, const T *rn,        \
                                      const T *cn, const T *P, int *out, int none);              \
  hipError_t launch_gather_rows(hipStream_t s, long n, int d, const T *X, const int *idx, T *out);
KM_TYPED_LAUNCHERS(float)
KM_TYPED_LAUNCHERS(double)
#undef KM_TYPED_LAUNCHERS
// rows (nullable): point indices of the m rows; out_pos (nullable): output
// slot of each row (default: rows[p], else p); few_rows: the rows are a short
// list (64-point workgroups instead of 128); kbest (nullable, few_rows and
// no best only): split_rows_cap() zeroed keys that let a list of at most that
// many rows be searched with the centres split over workgroups (left zeroed
// afterwards)
hipError_t launch_assign_f32(hipStream_t s, long m, const int *m_dev, int d, const float *X,
                             const int *rows, int k, const float *Ct, int ldc, const float *nc,
                             int *labels, float *best, const int *out_pos = nullptr,
                             bool few_rows = false, unsigned long long *kbest = nullptr);
long split_rows_cap();
hipError_t launch_assign_f64(hipStream_t s, long m, int d, const double *X, int k,
                             const double *Ct, int ldc, const double *nc, int *labels);
hipError_t launch_comp_reduce(hipStream_t s, int k, const double *compj, double *out);
hipError_t launch_update_zero(hipStream_t s, int d, int k, double *S, long long *cnt, int *empty);
hipError_t launch_generate(hipStream_t s, uint64_t seed, long i0, long n, int d, int ncomp,
                           const float *means, const float *sigmas, const double *cdf, float *X,
                           int *z);
int centers_ld(int k);
int num_cus();

// Distances from scores: |x - c|^2 = |x|^2 - 2 (x.c - 0.5 |c|^2), with xn >=
// |x| the point's norm bound.  dist_upper: an upper bound of |x - c| given lo
// <= the exact score of c; dist_lower: a lower bound given hi >= it.  (|x|^2
// lies between xn^2 (1 - 1.8e-7 D - 4e-6) and xn^2 (1 + 2.4e-7); the relative
// terms cover the roundings of the subtraction and of the root.)
__device__ inline float dist_upper(float xn, float lo) {
  const float x2u = xn * xn * (1.0f + 2.4e-7f);
  const float uu = (x2u - 2.0f * lo) + 3.6e-7f * (x2u + 2.0f * fabsf(lo));
  return sqrtf(fmaxf(uu, 0.0f)) * (1.0f + 1e-6f) + 1e-30f;
}
__device__ inline float dist_lower(float xn, float hi, float Df) {
  const float x2l = xn * xn * (1.0f - Df * 1.8e-7f - 4e-6f);
  const float ll = (x2l - 2.0f * hi) - 3.6e-7f * (x2l + 2.0f * fabsf(hi));
  return ll > 0.0f ? sqrtf(ll) * (1.0f - 1e-6f) : 0.0f;
}

// Upper bound of the distance a centre moved, from its FP32 coordinate
// differences t_q: the chain of their squares with its rounding covered
// (gamma_D in any order) when the largest |t_q| is at least 2^-40, else
// sqrt(D) max |t_q| (squares of smaller differences may underflow); 0
// exactly when every difference is 0, NaN when one is.
__device__ inline float drift_bound(float q, float ma, int D) {
  if (ma == 0.0f) return 0.0f;
  return ma >= 9.094947e-13f  // 2^-40
             ? sqrtf(q * (1.0f + (float)D * 1.2e-7f + 4e-7f)) * (1.0f + 1e-6f)
             : sqrtf((float)D) * ma * (1.0f + 1e-6f) + 1e-45f;
}

}  // namespace km
