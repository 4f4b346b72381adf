// Plain FP32 path for gfx950: fused score+argmax, canonical FP64 update,
// compactness, norms, synthetic-data generator.
//
// Reference counterparts (paths relative to the reference repo):
//   k_assign_f32_mfma   main/cuda/ClosestCenters.cu:46-123 (+ multiply.hcu,
//                       dotproducts.hcu, fma.hcu, reduction.hcu, vectorMin.hcu)
//   update kernels      main/cuda/ClusterCenters.cu:3-132
//   compactness         main/cuda/Compactness.cu:3-96
//   k_col_norms         main/cuda/Normalize.cu:33-47
//   k_row_norms         main/cuda/Normalize.cu:3-31
#include <float.h>
#include <limits.h>
#include <math.h>

#include "km_common.hpp"

namespace km {

int centers_ld(int k) { return (int)(km_cdiv(k, 256) * 256); }

int num_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

// Element types: float (the FP32 engine) and double (the FP64 engine,
// computeKmeansD and the ...D kernel names).  Every chain below is written
// with explicit fused multiply-adds in the element type.
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------- norms
// nc[j] = fma chain over d of C[j][d]^2 (Normalize.cu:41-44, contraction explicit).
template <typename T>
__global__ void k_col_norms(int k, int d, const T *__restrict__ C, T *__restrict__ nc) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const T *c = C + (long)j * d;
  T s = 0;
  for (int q = 0; q < d; ++q) s = fma_t(c[q], c[q], s);
  nc[j] = s;
}

// Row norms: the same chain per row (Normalize.cu:3-31).  256 rows per
// block; slices of 16 dimensions are read row-contiguously into LDS (16
// lanes per row) and each thread then runs its row's chain from LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_row_norms(long m, int d, const T *__restrict__ X,
                                                   T *__restrict__ out) {
  __shared__ T tile[256][17];
  const long r0 = (long)blockIdx.x * 256;
  const int rows = (int)min(256L, m - r0);
  T s = 0;
  for (int d0 = 0; d0 < d; d0 += 16) {
    const int dn = min(16, d - d0);
    __syncthreads();
    for (int e = threadIdx.x; e < rows * 16; e += 256) {
      const int r = e >> 4, q = e & 15;
      if (q < dn) tile[r][q] = X[(r0 + r) * d + d0 + q];
    }
    __syncthreads();
    if ((int)threadIdx.x < rows)
      for (int q = 0; q < dn; ++q) s = fma_t(tile[threadIdx.x][q], tile[threadIdx.x][q], s);
  }
  if ((int)threadIdx.x < rows) out[r0 + threadIdx.x] = s;
}

// Ct[q][j] = C[j][q], zero for j >= k (Ct pre-cleared by the caller once).
template <typename T>
__global__ void k_transpose(int k, int d, const T *__restrict__ C, T *__restrict__ Ct, int ldc) {
  __shared__ T t[32][33];
  int j0 = blockIdx.x * 32, q0 = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    int j = j0 + r, q = q0 + tx;
    t[r][tx] = (j < k && q < d) ? C[(long)j * d + q] : (T)0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int q = q0 + r, j = j0 + tx;
    if (q < d && j < ldc) Ct[(long)q * ldc + j] = t[tx][r];
  }
}

// ---------------------------------------------------------------- K1 assign, matrix cores
// v_mfma_f32_16x16x4f32 and v_mfma_f64_16x16x4f64 compute D[i][j] = C[i][j] +
// sum_k A[i][k] B[k][j] as four fused multiply-adds in k order, each rounded
// once: measured on gfx950 against the sequential fma chain, bit for bit,
// including subnormal, infinite, NaN and signed-zero operands.  Chaining the
// instruction over the dimensions in steps of four, from C = -0.5*nc[j], is
// therefore the canonical score chain itself, at the matrix cores' rate (for
// FP32 twice the unpacked VALU rate; for FP64 the same rate, without the
// VALU kernel's LDS traffic).  Dimensions past D are padded with x = +0,
// c = -0: the product -0 leaves every running value unchanged (also -0).
//
// Layout of one 16x16x4 step (lane l, g = l / 16, r = l % 16): A = point r,
// dimension g; B = dimension g, centre r; result register v = point 4g + v
// (FP32) or 4v + g (FP64), centre r.  A wave owns 16*RT points with DS
// dimensions of them in registers; the workgroup's four waves sweep the
// centres in blocks of 64 (four column tiles), staged through a double
// buffer in LDS so that a lane's four B operands of a 16-dimension group are
// one read, and the next block is fetched one chunk per 16 dimensions while
// the current block's MFMAs run.
template <typename T> struct MfmaCfg;
template <> struct MfmaCfg<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static constexpr int RT = 2, DS = 128;
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};
template <> struct MfmaCfg<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static constexpr int RT = 2, DS = 128;
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return 4 * v + g; }
};
#define AM_CT 4
// points per workgroup (RT row tiles of 16 per wave)
template <typename T, int RT = MfmaCfg<T>::RT> constexpr int am_bm() { return 4 * 16 * RT; }

// Order-preserving 64-bit key of (score, centre): a larger key is a larger
// score, or the lower centre index on equal scores, so the maximum key over
// any partition of the centres is the single pass's (score, label).  Key 0
// lies below every key of a non-NaN score.  -0 and +0 compare equal in the
// single pass (a later centre must be strictly greater), so both get the
// bits of +0 here: otherwise a +0 in a later slice would beat a -0 at a lower
// index.  NaN never reaches the key: each slice's best is taken with '>'.
__device__ __forceinline__ unsigned long long split_key(float v, int j) {
  const unsigned u = v == 0.0f ? 0u : __float_as_uint(v);
  const unsigned o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)o << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)j);
}

// SPLIT: the work items are (tile, slice of the centre blocks); each item's
// best (score, centre) per point goes into kbest[point position] by a 64-bit
// atomic maximum of split_key, and k_split_labels writes the labels.  For
// short row lists, where whole tiles would leave most CUs idle.
template <typename T, int RT, bool SPLIT = false>
__device__ __forceinline__ void assign_mfma_tiles(long M, const int *__restrict__ Mdev, int D,
                                                  const T *__restrict__ X,
                                                  const int *__restrict__ rows, int K,
                                                  const T *__restrict__ Ct, int ldc,
                                                  const T *__restrict__ nc,
                                                  const int *__restrict__ out_pos,
                                                  int *__restrict__ labels, T *__restrict__ best,
                                                  long mlo, long mhi,
                                                  unsigned long long *__restrict__ kbest = nullptr) {
  typedef MfmaCfg<T> Cfg;
  typedef typename Cfg::v4 v4;
  constexpr int DS = Cfg::DS, BM = am_bm<T, RT>();
  // Bl[b][s][g][c] = the four dimensions 16 s + 4 t + g (t = 0..3) of column c,
  // two buffers
  __shared__ v4 Bl[2][DS / 16][4][64];
  if (Mdev) M = *Mdev;
  if (M < mlo || M > mhi) return;  // the row count is served by the other tile shape
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const long ntiles = (M + BM - 1) / BM;
  const int nblk = (K + 63) / 64;
  const int nsl = (D + DS - 1) / DS;
  // centre-block slices per tile: as many items as workgroups, at most
  long nslice = 1;
  if constexpr (SPLIT) {
    if (ntiles == 0) return;
    nslice = (long)gridDim.x / ntiles;
    if (nslice > nblk) nslice = nblk;
    if (nslice < 1) nslice = 1;
  }
  const long nitems = ntiles * nslice;
  for (long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const long tile = item / nslice, slice = item - tile * nslice;
    const int jb0 = (int)((long)nblk * slice / nslice);
    const int jb1 = (int)((long)nblk * (slice + 1) / nslice);
    const long p0 = tile * BM + w * (16 * RT);
    long xrow[RT];
    bool xok[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const long p = p0 + rt * 16 + r;
      xok[rt] = p < M;
      xrow[rt] = xok[rt] ? (rows ? (long)rows[p] : p) : 0;
    }
    T bv[RT][4];
    int bi[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) { bv[rt][v] = (T)-INFINITY; bi[rt][v] = 0; }
    T xa[RT][DS / 4];
    // chunk i of a staged block for thread (w, c): dimensions 16 i + 4 t + w
    // (t = 0..3) of column c, i.e. Bl[.][i][w][c]
    auto fetch = [&](int i, int jbn, int sln) {
      const int d0n = sln * DS, dnn = min(DS, D - d0n), c = threadIdx.x & 63;
      const T *src = Ct + (long)(d0n + 16 * i + w) * ldc + jbn * 64 + c;
      v4 v;
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t] = 16 * i + 4 * t + w < dnn ? src[(long)4 * t * ldc] : (T)-0.0;
      return v;
    };
    int buf = 0;
#pragma unroll
    for (int i = 0; i < DS / 16; ++i) Bl[0][i][w][threadIdx.x & 63] = fetch(i, jb0, 0);
    __syncthreads();
    for (int jb = jb0; jb < jb1; ++jb) {
      v4 acc[RT][AM_CT];
#pragma unroll
      for (int ct = 0; ct < AM_CT; ++ct) {
        const int j = jb * 64 + ct * 16 + r;
        const T v = j < K ? (T)-0.5 * nc[j] : (T)-INFINITY;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[rt][ct] = (v4){v, v, v, v};
      }
      for (int sl = 0; sl < nsl; ++sl) {
        const int d0 = sl * DS, dn = min(DS, D - d0);
        if (nsl > 1 || jb == jb0) {
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int q = 0; q < DS / 4; ++q) {
              const int d = d0 + 4 * q + g;
              xa[rt][q] = (xok[rt] && 4 * q + g < dn) ? X[xrow[rt] * D + d] : (T)0;
            }
        }
        // the next block is fetched one chunk per 16 dimensions into the other
        // buffer, so its loads are in flight during this block's MFMAs
        const bool nxt = sl + 1 < nsl || jb + 1 < jb1;
        const int jbn = sl + 1 < nsl ? jb : jb + 1, sln = sl + 1 < nsl ? sl + 1 : 0;
        const int ns = (dn + 15) / 16;
        v4 fl;
#pragma unroll
        for (int s16 = 0; s16 < DS / 16; ++s16) {  // unrolled: xa stays in registers
          if (nxt) {
            if (s16 > 0) Bl[buf ^ 1][s16 - 1][w][threadIdx.x & 63] = fl;
            fl = fetch(s16, jbn, sln);
          }
          if (s16 < ns) {
            v4 bf[AM_CT];
#pragma unroll
            for (int ct = 0; ct < AM_CT; ++ct) bf[ct] = Bl[buf][s16][g][ct * 16 + r];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
              for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int ct = 0; ct < AM_CT; ++ct)
                  acc[rt][ct] = Cfg::mfma(xa[rt][4 * s16 + t], bf[ct][t], acc[rt][ct]);
          }
        }
        if (nxt) Bl[buf ^ 1][DS / 16 - 1][w][threadIdx.x & 63] = fl;
        __syncthreads();  // next buffer written, this one no longer read
        buf ^= 1;
      }
      // running best per (point, lane); centres visited in ascending order
#pragma unroll
      for (int ct = 0; ct < AM_CT; ++ct) {
        const int j = jb * 64 + ct * 16 + r;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (acc[rt][ct][v] > bv[rt][v]) { bv[rt][v] = acc[rt][ct][v]; bi[rt][v] = j; }
      }
    }
    // the 16 lanes of a row group hold the same points; merge with the tie rule
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
          const T ov = __shfl_xor(bv[rt][v], off);
          const int oi = __shfl_xor(bi[rt][v], off);
          if (ov > bv[rt][v] || (ov == bv[rt][v] && oi < bi[rt][v])) { bv[rt][v] = ov; bi[rt][v] = oi; }
        }
    if (r == 0) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long pt = p0 + rt * 16 + Cfg::row(g, v);
          if (pt < M) {
            if constexpr (SPLIT) {
              atomicMax(kbest + pt, split_key((float)bv[rt][v], bi[rt][v]));
            } else {
              const long o = out_pos ? (long)out_pos[pt] : rows ? (long)rows[pt] : pt;
              labels[o] = bi[rt][v];
              if (best) best[o] = bv[rt][v];
            }
          }
        }
    }
  }
}

// RT = 1 (64 points per workgroup) serves short lists of rows (the points
// a reduced-precision pass leaves open), where 128-point tiles leave most
// CUs idle.  A kernel only works when the row count lies in [mlo, mhi]: for a
// count known only on the device both shapes are launched with complementary
// ranges and one of them exits at once.
template <int RT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_assign_f32_mfma(long M, const int *__restrict__ Mdev, int D, const float *__restrict__ X,
                       const int *__restrict__ rows, int K, const float *__restrict__ Ct, int ldc,
                       const float *__restrict__ nc, const int *__restrict__ out_pos,
                       int *__restrict__ labels, float *__restrict__ best, long mlo, long mhi) {
  assign_mfma_tiles<float, RT>(M, Mdev, D, X, rows, K, Ct, ldc, nc, out_pos, labels, best, mlo,
                               mhi);
}

// 64-point tiles with the centres split over workgroups (see SPLIT above)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_assign_f32_mfma_split(long M, const int *__restrict__ Mdev, int D,
                             const float *__restrict__ X, const int *__restrict__ rows, int K,
                             const float *__restrict__ Ct, int ldc, const float *__restrict__ nc,
                             long mlo, long mhi, unsigned long long *__restrict__ kbest) {
  assign_mfma_tiles<float, 1, true>(M, Mdev, D, X, rows, K, Ct, ldc, nc, nullptr, nullptr,
                                    nullptr, mlo, mhi, kbest);
}

// Labels from the split search's keys; each key is reset to 0 for the next
// search.
__global__ void k_split_labels(long M, const int *__restrict__ Mdev, long mlo, long mhi,
                               const int *__restrict__ rows, const int *__restrict__ out_pos,
                               unsigned long long *__restrict__ kbest, int *__restrict__ labels) {
  if (Mdev) M = *Mdev;
  if (M < mlo || M > mhi) return;
  for (long pt = (long)blockIdx.x * blockDim.x + threadIdx.x; pt < M;
       pt += (long)gridDim.x * blockDim.x) {
    const unsigned long long key = kbest[pt];
    kbest[pt] = 0ull;
    const long o = out_pos ? (long)out_pos[pt] : rows ? (long)rows[pt] : pt;
    labels[o] = (int)(0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFull));
  }
}

// one wave per SIMD: the double operands of 128 dimensions stay in registers
// (accumulators in AGPRs) and the double buffer takes 128 KB of LDS; measured
// 10% faster than two waves with 64-dimension slices re-read per block
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
void k_assign_f64_mfma(long M, const int *__restrict__ Mdev, int D, const double *__restrict__ X,
                       const int *__restrict__ rows, int K, const double *__restrict__ Ct, int ldc,
                       const double *__restrict__ nc, const int *__restrict__ out_pos,
                       int *__restrict__ labels, double *__restrict__ best) {
  assign_mfma_tiles<double, MfmaCfg<double>::RT>(M, Mdev, D, X, rows, K, Ct, ldc, nc, out_pos,
                                                 labels, best, 0L, LONG_MAX);
}


// ---------------------------------------------------------------- update
// Sub-block label histogram, label-major: hist[j*nsb + s].
__global__ void k_hist(const int *__restrict__ labels, long nb, int K, int nsb, int SB,
                       int *__restrict__ hist) {
  extern __shared__ int h[];
  for (int j = threadIdx.x; j < K; j += blockDim.x) h[j] = 0;
  __syncthreads();
  const int s = blockIdx.x;
  const long a = (long)s * SB, e = min(a + SB, nb);
  // 16 labels per thread in flight before their counts are added
  for (long b0 = a; b0 < e; b0 += 16L * blockDim.x) {
    int v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const long i = b0 + (long)u * blockDim.x + threadIdx.x;
      v[u] = i < e ? labels[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (v[u] >= 0) atomicAdd(&h[v[u]], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < K; j += blockDim.x) hist[(long)j * nsb + s] = h[j];
}

// Block-wide exclusive scan of one int per thread (256 threads).
__device__ inline int block_excl_scan256(int v, int *lds, int *total) {
  const int tid = threadIdx.x;
  lds[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int t = tid >= off ? lds[tid - off] : 0;
    __syncthreads();
    lds[tid] += t;
    __syncthreads();
  }
  int incl = lds[tid];
  if (total) *total = lds[255];
  __syncthreads();
  return incl - v;
}

#define SCAN_TILE 4096
__global__ __launch_bounds__(256) void k_scan_reduce(const int *__restrict__ in, long L,
                                                     int *__restrict__ tsum) {
  __shared__ int lds[256];
  const long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * 16;
  int s = 0;
  for (int q = 0; q < 16; ++q) s += (base + q < L) ? in[base + q] : 0;
  int tot;
  block_excl_scan256(s, lds, &tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_tiles(int *__restrict__ tsum, int T) {
  __shared__ int lds[256];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b = 0; b < T; b += 256) {
    int v = (b + threadIdx.x < T) ? tsum[b + threadIdx.x] : 0;
    int tot;
    int ex = block_excl_scan256(v, lds, &tot);
    if (b + threadIdx.x < T) tsum[b + threadIdx.x] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_scan_down(const int *__restrict__ in, long L,
                                                   const int *__restrict__ tsum,
                                                   int *__restrict__ out) {
  __shared__ int lds[256];
  const long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * 16;
  int v[16];
  int s = 0;
  for (int q = 0; q < 16; ++q) { v[q] = (base + q < L) ? in[base + q] : 0; s += v[q]; }
  int tot;
  int ex = block_excl_scan256(s, lds, &tot) + tsum[blockIdx.x];
  for (int q = 0; q < 16; ++q) {
    if (base + q < L) out[base + q] = ex;
    ex += v[q];
  }
  if (base <= L - 1 && base + 16 >= L) out[L] = ex;  // total after the last element
}

// Stable scatter of one sub-block of W * SBW points (blockDim = 64 W):
// wave w walks points [a + w SBW, a + (w + 1) SBW) in order.  The waves
// first count their labels into W x K LDS counters, which become each
// wave's write positions: the sub-block's offsets (offs) plus the counts of
// the waves before it.  Lanes with equal labels are then ranked by a match:
// with PEER, a per-wave LDS mask per label class (label mod 64, one 64-bit
// LDS OR per step) narrowed by a ballot per remaining label bit; otherwise a
// ballot per label bit.  The labels of 16 steps (1024 points) are read
// into registers at once and the next 16 are in flight while these are
// ranked.  A sub-block's W waves write one run per label, so neighbouring
// waves fill the same cache lines.
#define SC_STEPS 16
// the staged scatter: 8 waves, 16 steps of 64 rows per wave and tile
#define SCS_THREADS 512
#define SCS_STEPS 16
#define SCS_TILE (SCS_THREADS * SCS_STEPS)  // 8192 rows
#define SCS_SH 13                          // bits of a row within a tile
#define SCS_MAXK 1024
#define SCS_MIN_SUBBLOCKS 32  // below this the 16-wave scatter is faster
template <bool PEER>
__global__ __launch_bounds__(1024) void k_scatter(const int *__restrict__ labels, long i0, long nb,
                                                  int K, int nsb, int SB, int SBW, int nbits,
                                                  const int *__restrict__ offs,
                                                  int *__restrict__ perm) {
  extern __shared__ __attribute__((aligned(8))) int hw[];  // W x K counters, then W x 64 masks
  const int W = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned long long *pm = reinterpret_cast<unsigned long long *>(hw + ((W * K + 1) & ~1)) + w * 64;
  for (int t = threadIdx.x; t < W * K; t += blockDim.x) hw[t] = 0;
  if (PEER) pm[lane] = 0ull;
  __syncthreads();
  const long a = (long)blockIdx.x * SB + (long)w * SBW, e = min(a + SBW, nb);
  for (long b0 = a; b0 < e; b0 += 64 * SC_STEPS) {  // SC_STEPS labels per lane in flight
    int v[SC_STEPS];
#pragma unroll
    for (int u = 0; u < SC_STEPS; ++u) {
      const long i = b0 + u * 64 + lane;
      v[u] = i < e ? labels[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < SC_STEPS; ++u)
      if (v[u] >= 0) atomicAdd(&hw[w * K + v[u]], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < K; j += blockDim.x) {
    int run = offs[(long)j * nsb + blockIdx.x];
    for (int u = 0; u < W; ++u) {
      const int t = hw[u * K + j];
      hw[u * K + j] = run;
      run += t;
    }
  }
  __syncthreads();
  int *cnt = hw + w * K;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int cur[SC_STEPS], nxt[SC_STEPS];
  auto fetch = [&](long b0, int (&v)[SC_STEPS]) {
#pragma unroll
    for (int u = 0; u < SC_STEPS; ++u) {
      const long i = b0 + u * 64 + lane;
      v[u] = i < e ? labels[i] : 0;
    }
  };
  fetch(a, cur);
  for (long b0 = a; b0 < e; b0 += 64 * SC_STEPS) {
    fetch(b0 + 64 * SC_STEPS, nxt);  // zeros past the wave's range
#pragma unroll
    for (int u = 0; u < SC_STEPS; ++u) {
      const long i = b0 + u * 64 + lane;
      const bool ok = i < e;
      const int l = cur[u];
      uint64_t peers, cls = 0ull;
      if (PEER) {  // lanes of the same label class, then the higher bits
        if (ok) atomicOr(&pm[l & 63], 1ull << lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        cls = ok ? pm[l & 63] : 0ull;
        peers = cls;
        for (int b = 6; b < nbits; ++b) {
          const bool bit = (l >> b) & 1;
          const uint64_t m = __ballot(bit);
          peers &= bit ? m : ~m;
        }
      } else {
        peers = __ballot(ok);
        for (int b = 0; b < nbits; ++b) {
          const bool bit = (l >> b) & 1;
          const uint64_t m = __ballot(bit);
          peers &= bit ? m : ~m;
        }
      }
      // the counters are the wave's own: its LDS reads and writes run in
      // program order, so ordering the compiler's schedule is enough, and
      // the perm stores are never waited for
      int c = 0;
      if (ok) {
        c = cnt[l];
        perm[c + __popcll(peers & below)] = (int)(i0 + i);
      }
      __builtin_amdgcn_wave_barrier();
      if (ok && (peers >> lane) == 1ull) cnt[l] = c + __popcll(peers);
      if (PEER && ok && (cls >> lane) == 1ull) pm[l & 63] = 0ull;  // the class's last lane
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int u = 0; u < SC_STEPS; ++u) cur[u] = nxt[u];
  }
}

// The same scatter with its writes staged in LDS (k <= SCS_MAXK): one
// SCS_THREADS-thread workgroup per sub-block, its rows in tiles of SCS_TILE
// consecutive rows.  Each wave ranks its SCS_STEPS x 64 rows of a tile among
// equal labels (per-wave LDS lane masks per label class, label mod 64, plus
// ballots on the higher bits; 16-bit per-(wave, label) counts), the counts
// are scanned over waves and labels, every row goes to its label-major slot
// of the tile in LDS, and the block writes the tile in slot order: a label's
// members of one tile are consecutive in perm, so the writes come in runs of
// ~SCS_TILE / k entries instead of one scattered 4-byte write per point.
// Each thread issues exactly SCS_TILE / SCS_THREADS stores per tile (slots
// past the sub-block's end write dummy[thread]), so the next tile's labels,
// loaded before them, are waited for without waiting for the stores.
__global__ __launch_bounds__(SCS_THREADS) void k_scatter_staged(
    const int *__restrict__ labels, long i0, long nb, int K, int nsb, int SB, int nbits,
    const int *__restrict__ offs, int *__restrict__ perm, int *__restrict__ dummy) {
  constexpr int NW = SCS_THREADS / 64;
  extern __shared__ __attribute__((aligned(16))) int ss[];
  int *stage = ss;                        // SCS_TILE: local row | label << SCS_SH
  // int arrays padded to even lengths: the 64-bit masks (LDS atomics) need
  // 8-byte alignment
  int *gc = stage + SCS_TILE;             // K: each label's next perm position
  int *toff = gc + ((K + 1) & ~1);        // K + 1: the tile's label offsets
  unsigned long long *pm =
      reinterpret_cast<unsigned long long *>(toff + ((K + 2) & ~1));  // NW x 64 class masks
  uint16_t *cnt = reinterpret_cast<uint16_t *>(pm + NW * 64);        // NW x K counts
  __shared__ int tmp[2 * NW];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const long a0 = (long)blockIdx.x * SB;
  const int rows_sb = (int)min((long)SB, nb - a0);  // > 0
  const int *lab = labels + a0;
  for (int j = t; j < K; j += SCS_THREADS) gc[j] = offs[(long)j * nsb + blockIdx.x];
  for (int e = t; e < NW * K; e += SCS_THREADS) cnt[e] = 0;
  for (int e = t; e < NW * 64; e += SCS_THREADS) pm[e] = 0ull;
  __syncthreads();
  uint16_t *cw = cnt + w * K;
  unsigned long long *pw = pm + w * 64;
  const int wr = w * (64 * SCS_STEPS);  // the wave's first row in a tile
  int Lc[SCS_STEPS], Ln[SCS_STEPS];
  auto fetch = [&](int t0, int (&v)[SCS_STEPS]) {  // clamped into the sub-block
#pragma unroll
    for (int u = 0; u < SCS_STEPS; ++u) v[u] = lab[min(t0 + wr + u * 64 + lane, rows_sb - 1)];
  };
  // labels clamped into [0, K) where they arrive (this also waits for the
  // first tile's loads here, outside the tile loop)
  fetch(0, Lc);
#pragma unroll
  for (int u = 0; u < SCS_STEPS; ++u) Lc[u] = min(max(Lc[u], 0), K - 1);
  for (int t0 = 0; t0 < rows_sb; t0 += SCS_TILE) {
    fetch(t0 + SCS_TILE, Ln);  // unconditional: the last row past the end
    // ranks of the wave's rows among its rows of equal label (row order)
    int lr[SCS_STEPS];
#pragma unroll
    for (int u = 0; u < SCS_STEPS; ++u) {
      const bool ok = t0 + wr + u * 64 + lane < rows_sb;
      const int L = Lc[u];
      if (ok) atomicOr(&pw[L & 63], 1ull << lane);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint64_t cls = ok ? pw[L & 63] : 0ull;
      uint64_t m = cls;
      for (int bt = 6; bt < nbits; ++bt) {
        const bool bit = (L >> bt) & 1;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
      }
      const int c = ok ? (int)cw[L] : 0;
      lr[u] = c + __popcll(m & below);
      __builtin_amdgcn_wave_barrier();
      if (ok && (m >> lane) == 1ull) cw[L] = (uint16_t)(c + __popcll(m));  // the label's last lane
      if (ok && (cls >> lane) == 1ull) pw[L & 63] = 0ull;                  // the class's last lane
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per label: exclusive prefix of the counts over waves (in place) and
    // the tile's label offsets (exclusive prefix over labels; each thread
    // owns labels 2t and 2t + 1)
    int s0 = 0, s1 = 0;
    if (2 * t < K) {
      for (int ww = 0; ww < NW; ++ww) {
        const int c = cnt[ww * K + 2 * t];
        cnt[ww * K + 2 * t] = (uint16_t)s0;
        s0 += c;
      }
    }
    if (2 * t + 1 < K) {
      for (int ww = 0; ww < NW; ++ww) {
        const int c = cnt[ww * K + 2 * t + 1];
        cnt[ww * K + 2 * t + 1] = (uint16_t)s1;
        s1 += c;
      }
    }
    const int s = s0 + s1;
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    if (t < 64) {
      const int wt = t < NW ? tmp[t] : 0;
      int z = wt;
#pragma unroll
      for (int o = 1; o < NW; o <<= 1) {
        const int y = __shfl_up(z, o);
        if (t >= o) z += y;
      }
      if (t < NW) tmp[NW + t] = z - wt;
    }
    __syncthreads();
    {
      const int base = tmp[NW + w] + x - s;
      if (2 * t < K) toff[2 * t] = base;
      if (2 * t + 1 < K) toff[2 * t + 1] = base + s0;
      if (t == SCS_THREADS - 1) toff[K] = base + s;  // the tile's rows
    }
    __syncthreads();
    // every row to its label-major slot of the tile
#pragma unroll
    for (int u = 0; u < SCS_STEPS; ++u) {
      const int r = wr + u * 64 + lane;
      const int L = Lc[u];
      if (t0 + r < rows_sb) stage[toff[L] + cw[L] + lr[u]] = r | (L << SCS_SH);
    }
    __syncthreads();
    // the wave's counts back to zero for the next tile (only this wave
    // writes them before the next tile's first barrier)
    for (int j = lane; j < K; j += 64) cw[j] = 0;
    // the tile in slot order: runs of one label go to consecutive positions
    const int rows = min(SCS_TILE, rows_sb - t0);
    const int rb = (int)(i0 + a0) + t0;  // the block's rows are below 2^31
#pragma unroll
    for (int q = 0; q < SCS_TILE / SCS_THREADS; ++q) {
      const int i = t + q * SCS_THREADS;
      const bool ok = i < rows;
      const int v = stage[ok ? i : 0];
      const int L = v >> SCS_SH;
      int *dst = ok ? perm + gc[L] + (i - toff[L]) : dummy + t;
      *dst = rb + (v & (SCS_TILE - 1));
    }
    __syncthreads();
    for (int j = t; j < K; j += SCS_THREADS) gc[j] += toff[j + 1] - toff[j];
#pragma unroll
    for (int u = 0; u < SCS_STEPS; ++u) Lc[u] = min(max(Ln[u], 0), K - 1);
  }
}

// Exclusive scan of one int per thread over a 1024-thread block.
__device__ inline int block_excl_scan1024(int v, int *wsum, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int t = wsum[u];
    if (u < w) pre += t;
    tot += t;
  }
  *total = tot;
  return pre + x - v;
}

// Per-centre member range in the sorted order and chunk prefix: one block
// of 1024 threads, each owning ceil(K/1024) consecutive centres, one scan.
__global__ __launch_bounds__(1024) void k_cluster_setup(int K, int nsb, const int *__restrict__ offs,
                                                        int *__restrict__ cstart,
                                                        int *__restrict__ ccount,
                                                        int *__restrict__ cq) {
  __shared__ int wsum[16];
  const int R = (K + 1023) / 1024;
  const int j0 = (int)threadIdx.x * R, j1 = min(K, j0 + R);
  int q = 0;
  for (int j = j0; j < j1; ++j) {
    const int st = offs[(long)j * nsb], n = offs[(long)(j + 1) * nsb] - st;
    cstart[j] = st;
    ccount[j] = n;
    q += (n + KM_CHUNK - 1) / KM_CHUNK;
  }
  int tot;
  int ex = block_excl_scan1024(q, wsum, &tot);
  for (int j = j0; j < j1; ++j) {
    cq[j] = ex;
    ex += (offs[(long)(j + 1) * nsb] - offs[(long)j * nsb] + KM_CHUNK - 1) / KM_CHUNK;
  }
  if (threadIdx.x == 0) cq[K] = tot;
}

// Read-once loads for the gathered rows (nontemporal: the update reads
// each row once, so the caches keep the sorted indices instead).
typedef float km_f4v __attribute__((ext_vector_type(4)));
__device__ inline float4 load_once(const float4 *p) {
  const km_f4v v = __builtin_nontemporal_load(reinterpret_cast<const km_f4v *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
template <typename T>
__device__ inline T load_once(const T *p) { return __builtin_nontemporal_load(p); }

// s + v * v.  For a float widened to double the square is exact, so the
// fused form, which rounds once, equals the unfused sum; a double square
// is rounded on its own first, as the oracle does.
template <typename T>
__device__ inline double add_square(double s, double v) {
  if (sizeof(T) == 4) return __builtin_fma(v, v, s);
  return s + v * v;
}

// Finds the centre of chunk c (the last j with cq[j] <= c) and the chunk's
// member range [start, start + cnt) in the sorted order.
__device__ inline void chunk_range(long c, int K, const int *__restrict__ cstart,
                                   const int *__restrict__ ccount, const int *__restrict__ cq,
                                   int &start, int &cnt, int &j, int &local) {
  int lo = 0, hi = K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cq[mid] <= c) lo = mid; else hi = mid - 1;
  }
  j = lo;
  local = (int)(c - cq[lo]);
  start = cstart[lo] + local * KM_CHUNK;
  cnt = min(KM_CHUNK, ccount[lo] - local * KM_CHUNK);
}

// Whether chunk `local` of centre j has the same members as the previous
// sort's chunk `local` of j (fd: j's first differing member position, cs_prev
// the previous [cstart | ccount | cq]), and if so that chunk's index there.
__device__ inline long chunk_reuse(int j, int local, int cnt_new, int K, const int *__restrict__ fd,
                                   const int *__restrict__ cs_prev) {
  if (!fd) return -1;
  const int q0 = local * KM_CHUNK;
  const int end_new = q0 + cnt_new;
  const int end_old = min(q0 + KM_CHUNK, cs_prev[K + j]);
  if (end_new != end_old || end_new > fd[j] || end_new <= q0) return -1;
  return (long)cs_prev[2 * K + j] + local;
}

// First-level sums: one group of Dp threads per chunk of <= KM_CHUNK members,
// one dimension per thread (slices of Dp dimensions), members added in
// ascending point order.  Each thread also sums the squares of its
// dimension (exact products for float rows, rounded ones for double rows,
// never fused); column D of the partial is those per-dimension sums added
// in ascending dimension order, so the squared norms come from the rows
// already read.  Every thread reaches every barrier: the slice loop is
// uniform over the block.
// With fd (nullable), chunks whose members the previous sort had in the same
// chunk copy that chunk's partials from P_prev instead (chunk_reuse).
template <typename T>
__global__ __launch_bounds__(256) void k_chunk_sums(const T *__restrict__ X, int D, int Dp,
                                                    const int *__restrict__ perm, int K,
                                                    const int *__restrict__ cstart,
                                                    const int *__restrict__ ccount,
                                                    const int *__restrict__ cq,
                                                    double *__restrict__ P,
                                                    const int *__restrict__ fd,
                                                    const int *__restrict__ cs_prev,
                                                    const double *__restrict__ P_prev) {
  __shared__ int pidx[4][KM_CHUNK];
  __shared__ double sqs[256];
  const int G = min(256 / Dp, 64);  // chunks per block
  const int sub = threadIdx.x / Dp;
  const int t = threadIdx.x - sub * Dp;
  const long c = (long)blockIdx.x * G + sub;
  const bool act = sub < G && c < cq[K];
  const int ld = D + 1;
  int cnt = 0, start = 0, j = 0, local = 0;
  long src = -1;
  if (act) {
    chunk_range(c, K, cstart, ccount, cq, start, cnt, j, local);
    src = chunk_reuse(j, local, cnt, K, fd, cs_prev);
  }
  const bool cp = src >= 0;  // copied, not computed
  // stage member indices (G <= 4 when Dp >= 64; otherwise read perm directly)
  const bool staged = G <= 4;
  if (act && !cp && staged)
    for (int m = t; m < cnt; m += Dp) pidx[sub][m] = perm[start + m];
  __syncthreads();
  double qs = 0.0;
  for (int d0 = 0; d0 < D; d0 += Dp) {
    const int d = d0 + t;
    double acc = 0.0, sq = 0.0;
    if (act && cp && d < D) P[c * ld + d] = P_prev[src * ld + d];
    if (act && !cp && d < D) {
      int m = 0;
      for (; m + 4 <= cnt; m += 4) {
        const int p0 = staged ? pidx[sub][m] : perm[start + m];
        const int p1 = staged ? pidx[sub][m + 1] : perm[start + m + 1];
        const int p2 = staged ? pidx[sub][m + 2] : perm[start + m + 2];
        const int p3 = staged ? pidx[sub][m + 3] : perm[start + m + 3];
        const double v0 = (double)load_once(X + (long)p0 * D + d);
        const double v1 = (double)load_once(X + (long)p1 * D + d);
        const double v2 = (double)load_once(X + (long)p2 * D + d);
        const double v3 = (double)load_once(X + (long)p3 * D + d);
        acc += v0; sq = add_square<T>(sq, v0);
        acc += v1; sq = add_square<T>(sq, v1);
        acc += v2; sq = add_square<T>(sq, v2);
        acc += v3; sq = add_square<T>(sq, v3);
      }
      for (; m < cnt; ++m) {
        const double v =
            (double)load_once(X + (long)(staged ? pidx[sub][m] : perm[start + m]) * D + d);
        acc += v;
        sq = add_square<T>(sq, v);
      }
      P[c * ld + d] = acc;
    }
    sqs[threadIdx.x] = sq;
    __syncthreads();
    if (act && !cp && t == 0) {
      const int dn = min(Dp, D - d0);
      for (int u = 0; u < dn; ++u) qs += sqs[sub * Dp + u];
    }
    __syncthreads();
  }
  if (act && t == 0) P[c * ld + D] = cp ? P_prev[src * ld + D] : qs;
}

// The same first-level sums with four dimensions per thread (D % 4 == 0,
// 16 <= D <= 256, rows 16-byte aligned and read as float4): TPC threads per
// chunk (D/4 rounded up to a power of two) and G = min(256/TPC, 64) chunks
// per block (G * D <= 1024); every dimension still adds its members in
// ascending point order, eight rows in flight per thread.  The square sums
// go through LDS to the chunk's first thread, which adds them in dimension
// order into column D.
__global__ __launch_bounds__(256) void k_chunk_sums4(const float *__restrict__ X, int D, int TPC,
                                                     const int *__restrict__ perm, int K,
                                                     const int *__restrict__ cstart,
                                                     const int *__restrict__ ccount,
                                                     const int *__restrict__ cq,
                                                     double *__restrict__ P,
                                                     const int *__restrict__ fd,
                                                     const int *__restrict__ cs_prev,
                                                     const double *__restrict__ P_prev) {
  __shared__ int pidx[16][KM_CHUNK];
  __shared__ double sqs[1024];
  const int G = min(256 / TPC, 64);
  const int sub = threadIdx.x / TPC;
  const int t = threadIdx.x - sub * TPC;
  const long c = (long)blockIdx.x * G + sub;
  const bool act = sub < G && c < cq[K];
  const int ld = D + 1;
  int cnt = 0, start = 0, j = 0, local = 0;
  long src = -1;
  if (act) {
    chunk_range(c, K, cstart, ccount, cq, start, cnt, j, local);
    src = chunk_reuse(j, local, cnt, K, fd, cs_prev);
  }
  const bool cp = src >= 0;  // copied, not computed
  const bool staged = G <= 16;
  if (act && !cp && staged)
    for (int m = t; m < cnt; m += TPC) pidx[sub][m] = perm[start + m];
  __syncthreads();
  if (act && cp) {
    for (int d = t; d <= D; d += TPC) P[c * ld + d] = P_prev[src * ld + d];
  } else if (act && 4 * t < D) {
    const float4 *__restrict__ X4 = reinterpret_cast<const float4 *>(X);
    const long D4 = D / 4;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    auto add = [&](const float4 v) {
      const double x = v.x, y = v.y, z = v.z, w = v.w;
      a0 += x; s0 = add_square<float>(s0, x);
      a1 += y; s1 = add_square<float>(s1, y);
      a2 += z; s2 = add_square<float>(s2, z);
      a3 += w; s3 = add_square<float>(s3, w);
    };
    int m = 0;
    if (staged) {
      for (; m + 8 <= cnt; m += 8) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = load_once(X4 + (long)pidx[sub][m + u] * D4 + t);
#pragma unroll
        for (int u = 0; u < 8; ++u) add(v[u]);
      }
    } else {
      // member indices from HBM: the next eight are loaded while this
      // batch's rows are in flight, so each batch waits for memory once
      int pn[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) pn[u] = u < cnt ? perm[start + u] : 0;
      for (; m + 8 <= cnt; m += 8) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = load_once(X4 + (long)pn[u] * D4 + t);
#pragma unroll
        for (int u = 0; u < 8; ++u) pn[u] = m + 8 + u < cnt ? perm[start + m + 8 + u] : 0;
#pragma unroll
        for (int u = 0; u < 8; ++u) add(v[u]);
      }
    }
    for (; m < cnt; ++m) add(load_once(X4 + (long)(staged ? pidx[sub][m] : perm[start + m]) * D4 + t));
    double *o = P + c * ld + 4 * t;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
    o[3] = a3;
    double *sq = sqs + sub * D + 4 * t;
    sq[0] = s0;
    sq[1] = s1;
    sq[2] = s2;
    sq[3] = s3;
  }
  __syncthreads();
  if (act && !cp && t == 0) {
    double qs = 0.0;
    for (int u = 0; u < D; ++u) qs += sqs[sub * D + u];
    P[c * ld + D] = qs;
  }
}

// Reuse: fd[j] = the first position at which centre j's member list in this
// sort differs from the previous sort's (the shorter count when one is a
// prefix of the other).  One workgroup per centre walks both lists 2048
// positions at a time and stops at the first step with a difference, so a
// centre with an early change costs little; lists read coalesced.
__global__ __launch_bounds__(256) void k_first_diff(int K, const int *__restrict__ perm,
                                                    const int *__restrict__ cstart,
                                                    const int *__restrict__ ccount,
                                                    const int *__restrict__ perm_prev,
                                                    const int *__restrict__ cs_prev,
                                                    int *__restrict__ fd) {
  __shared__ int first;
  const int j = blockIdx.x, t = threadIdx.x;
  if (j >= K) return;
  const int lim = min(ccount[j], cs_prev[K + j]);
  const int *a = perm + cstart[j], *b = perm_prev + cs_prev[j];
  if (t == 0) first = lim;
  __syncthreads();
  for (int m0 = 0; m0 < lim; m0 += 2048) {  // 8 positions per thread per step, loads issued together
    int va[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = m0 + u * 256 + t;
      va[u] = m < lim ? a[m] : 0;
      vb[u] = m < lim ? b[m] : 0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (va[u] != vb[u]) {
        atomicMin(&first, m0 + u * 256 + t);
        break;
      }
    __syncthreads();
    if (first < lim) break;  // uniform: every thread reads it after the barrier
    __syncthreads();         // the next step's atomics only after every thread's read
  }
  if (t == 0) fd[j] = first;
}

// Per-centre CSUM over its chunk partials (D sums and the q column),
// added into the running block sums.
__global__ __launch_bounds__(256) void k_block_finalize(int D, int K, const int *__restrict__ ccount,
                                                        const int *__restrict__ cq,
                                                        const double *__restrict__ P,
                                                        double *__restrict__ S,
                                                        long long *__restrict__ cnt) {
  const int j = blockIdx.x;
  const int q0 = cq[j], nq = cq[j + 1] - q0;
  const int ld = D + 1;
  for (int d = threadIdx.x; d <= D; d += blockDim.x) {
    double b = 0.0;
    if (nq <= KM_CHUNK) {
      int t = 0;
      for (; t + 8 <= nq; t += 8) {  // eight loads in flight, added in order
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = P[(long)(q0 + t + u) * ld + d];
#pragma unroll
        for (int u = 0; u < 8; ++u) b += v[u];
      }
      for (; t < nq; ++t) b += P[(long)(q0 + t) * ld + d];
    } else {  // nq <= KM_CHUNK^2 since KM_BLOCK <= KM_CHUNK^3
      for (int g = 0; g < nq; g += KM_CHUNK) {
        double gs = 0.0;
        const int ge = min(nq, g + KM_CHUNK);
        for (int t = g; t < ge; ++t) gs += P[(long)(q0 + t) * ld + d];
        b += gs;
      }
    }
    S[(long)j * ld + d] += b;
  }
  if (threadIdx.x == 0) cnt[j] += ccount[j];
}

// Clears the FP64 sums, the counts and (nullable) the empty-cluster counter
// with stores, not a memset, so the launch can sit in a captured graph.
__global__ void k_zero_sums(long n, double *__restrict__ S, int K, long long *__restrict__ cnt,
                            int *__restrict__ empty) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) S[i] = 0.0;
  if (i < K) cnt[i] = 0;
  if (i == 0 && empty) *empty = 0;
}

// New centres, transposed copy and norms; empty centres keep the old value.
// compj[j] = (Q_j - 2 sum_d c_jd S_jd) + n_j sum_d c_jd^2, the compactness
// contribution of centre j against the new centre (sums left to right).
template <typename T>
__global__ __launch_bounds__(256) void k_final_centres(int D, int K, const double *__restrict__ S,
                                                       const long long *__restrict__ cnt,
                                                       const T *C_old, T *C_new,
                                                       T *__restrict__ Ct, int ldc,
                                                       T *__restrict__ nc,
                                                       long long *__restrict__ counts,
                                                       int *__restrict__ empty,
                                                       double *__restrict__ compj,
                                                       int stage_sums) {
  // the new centre (T) and, when stage_sums, its sums (FP64) for the
  // compactness terms
  extern __shared__ __attribute__((aligned(8))) unsigned char fc_smem[];
  double *ss = reinterpret_cast<double *>(fc_smem);
  T *cs = reinterpret_cast<T *>(ss + (stage_sums ? D : 0));
  const int j = blockIdx.x;
  const long long n = cnt[j];
  const int ld = D + 1;
  for (int d = threadIdx.x; d < D; d += blockDim.x) {
    const double sv = S[(long)j * ld + d];
    const T v = n > 0 ? (T)(sv / (double)n) : C_old[(long)j * D + d];
    cs[d] = v;
    if (stage_sums) ss[d] = sv;
    C_new[(long)j * D + d] = v;
    if (Ct) Ct[(long)d * ldc + j] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (nc) {
      T s = 0;
      for (int d = 0; d < D; ++d) s = fma_t(cs[d], cs[d], s);
      nc[j] = s;
    }
    if (compj) {
      double a = 0.0, b = 0.0;
      for (int d = 0; d < D; ++d) {
        const double c = (double)cs[d];
        a += c * (stage_sums ? ss[d] : S[(long)j * ld + d]);
        b += c * c;
      }
      compj[j] = (S[(long)j * ld + D] - 2.0 * a) + (double)n * b;
    }
    if (counts) counts[j] = n;
    if (empty && n == 0) atomicAdd(empty, 1);
  }
}

// comp_j terms for arbitrary centres C (k x D) from the moments.
template <typename T>
__global__ void k_comp_terms(int D, int K, const double *__restrict__ S,
                             const long long *__restrict__ cnt, const T *__restrict__ C,
                             double *__restrict__ compj) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const int ld = D + 1;
  double a = 0.0, b = 0.0;
  for (int d = 0; d < D; ++d) {
    const double c = (double)C[(long)j * D + d];
    a += c * S[(long)j * ld + d];
    b += c * c;
  }
  compj[j] = (S[(long)j * ld + D] - 2.0 * a) + (double)cnt[j] * b;
}

// out = CSUM of compj[0..K) (K <= KM_CHUNK^2).  The terms are staged in LDS
// 4096 at a time (all 256 threads loading), then 16 threads add one
// KM_CHUNK piece each in order, so the serial sums wait on LDS, not memory.
__global__ __launch_bounds__(256) void k_comp_reduce(const double *__restrict__ compj, int K,
                                                     double *__restrict__ out) {
  __shared__ double st[16 * KM_CHUNK];
  __shared__ double g[KM_CHUNK];
  const int ng = (K + KM_CHUNK - 1) / KM_CHUNK;
  for (int r0 = 0; r0 < K; r0 += 16 * KM_CHUNK) {
    const int rn = min(16 * KM_CHUNK, K - r0);
    for (int t = threadIdx.x; t < rn; t += blockDim.x) st[t] = compj[r0 + t];
    __syncthreads();
    const int q = threadIdx.x;
    if (q < 16 && q * KM_CHUNK < rn) {
      const int e = min(rn, (q + 1) * KM_CHUNK);
      double sum = 0.0;
      for (int t = q * KM_CHUNK; t < e; ++t) sum += st[t];
      g[r0 / KM_CHUNK + q] = sum;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (ng <= 1) {
      *out = ng == 1 ? g[0] : 0.0;
    } else {
      double sum = 0.0;
      for (int q = 0; q < ng; ++q) sum += g[q];
      *out = sum;
    }
  }
}

// q_i = left-to-right FP64 sum over d of (double)x_id^2 (squares exact).
__global__ void k_point_sq(long m, int d, const float *__restrict__ X, double *__restrict__ q) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const float *x = X + i * d;
  double s = 0.0;
  if ((d & 3) == 0) {
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
#pragma unroll 8
    for (int t = 0; t < d / 4; ++t) {
      const float4 v = x4[t];
      s += (double)v.x * (double)v.x;
      s += (double)v.y * (double)v.y;
      s += (double)v.z * (double)v.z;
      s += (double)v.w * (double)v.w;
    }
  } else {
    for (int t = 0; t < d; ++t) s += (double)x[t] * (double)x[t];
  }
  q[i] = s;
}

// FP64 points: q_i = left-to-right sum over d of the rounded squares
// x_id * x_id (no fused multiply-add, so the same value as the oracle's).
// 256 rows per block, 16-dimension slices read row-contiguously into LDS.
__global__ __launch_bounds__(256) void k_point_sq_f64(long m, int d, const double *__restrict__ X,
                                                      double *__restrict__ q) {
  __shared__ double tile[256][17];
  const long r0 = (long)blockIdx.x * 256;
  const int rows = (int)min(256L, m - r0);
  double s = 0.0;
  for (int d0 = 0; d0 < d; d0 += 16) {
    const int dn = min(16, d - d0);
    __syncthreads();
    for (int e = threadIdx.x; e < rows * 16; e += 256) {
      const int r = e >> 4, c = e & 15;
      if (c < dn) tile[r][c] = X[(r0 + r) * d + d0 + c];
    }
    __syncthreads();
    if ((int)threadIdx.x < rows)
      for (int c = 0; c < dn; ++c) s += tile[threadIdx.x][c] * tile[threadIdx.x][c];
  }
  if ((int)threadIdx.x < rows) q[r0 + threadIdx.x] = s;
}

// ---------------------------------------------------------------- generator
__global__ void k_generate(uint64_t seed, long i0, long n, int d, int ncomp,
                           const float *__restrict__ means, const float *__restrict__ sigmas,
                           const double *__restrict__ cdf, float *__restrict__ X,
                           int *__restrict__ z) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const long i = i0 + r;
  const double u = rng_unif(seed, 2, (uint64_t)i);
  int lo = 0, hi = ncomp - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (u <= cdf[mid]) hi = mid; else lo = mid + 1;
  }
  if (z) z[r] = lo;
  const float sg = sigmas[lo];
  for (int q = 0; q < d; ++q) {
    const uint64_t k = (uint64_t)i * (uint64_t)d + (uint64_t)q;
    const double u1 = rng_unif(seed, 3, 2 * k), u2 = rng_unif(seed, 3, 2 * k + 1);
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    X[r * d + q] = (float)((double)means[(long)lo * d + q] + (double)sg * g);
  }
}

template <typename T>
__global__ void k_gather_rows(long n, int d, const T *__restrict__ X, const int *__restrict__ idx,
                              T *__restrict__ out) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n * d;
       e += (long)gridDim.x * blockDim.x) {
    const long r = e / d;
    const int q = (int)(e - r * d);
    out[e] = X[(long)idx[r] * d + q];
  }
}

// ---------------------------------------------------------------- launchers
template <typename T>
static hipError_t launch_col_norms_t(hipStream_t s, int k, int d, const T *C, T *nc) {
  hipLaunchKernelGGL(k_col_norms<T>, dim3(km_cdiv(k, 256)), dim3(256), 0, s, k, d, C, nc);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_row_norms_t(hipStream_t s, long m, int d, const T *X, T *out) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_norms<T>, dim3(km_cdiv(m, 256)), dim3(256), 0, s, m, d, X, out);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_transpose_t(hipStream_t s, int k, int d, const T *C, T *Ct, int ldc) {
  dim3 grid(km_cdiv(ldc, 32), km_cdiv(d, 32));
  hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, s, k, d, C, Ct, ldc);
  return hipGetLastError();
}

long split_rows_cap() { return (long)am_bm<float, 1>() * 2 * num_cus(); }

hipError_t launch_assign_f32(hipStream_t s, long m, const int *m_dev, int d, const float *X,
                             const int *rows, int k, const float *Ct, int ldc, const float *nc,
                             int *labels, float *best, const int *out_pos, bool few_rows,
                             unsigned long long *kbest) {
  const long cap = split_rows_cap();  // two 64-point tiles per CU
  const long ncap = (long)num_cus() * 8;
  if (m <= 0) return hipSuccess;
  if (few_rows && (m_dev != nullptr || m <= cap)) {
    const long t1 = km_cdiv(m < cap ? m : cap, am_bm<float, 1>());
    if (kbest && !best) {  // one item per workgroup slot: tiles times centre slices
      const long gs = 2L * num_cus();
      hipLaunchKernelGGL(k_assign_f32_mfma_split, dim3(gs), dim3(256), 0, s, m, m_dev, d, X,
                         rows, k, Ct, ldc, nc, 0L, cap, kbest);
      hipLaunchKernelGGL(k_split_labels, dim3(km_cdiv(m < cap ? m : cap, 256)), dim3(256), 0, s,
                         m, m_dev, 0L, cap, rows, out_pos, kbest, labels);
    } else {
      hipLaunchKernelGGL(k_assign_f32_mfma<1>, dim3(t1 < ncap ? t1 : ncap), dim3(256), 0, s, m,
                         m_dev, d, X, rows, k, Ct, ldc, nc, out_pos, labels, best, 0L, cap);
    }
    if (m_dev == nullptr && m <= cap) return hipGetLastError();
  }
  const long t2 = km_cdiv(m, am_bm<float, 2>());
  hipLaunchKernelGGL(k_assign_f32_mfma<2>, dim3(t2 < ncap ? t2 : ncap), dim3(256), 0, s, m, m_dev,
                     d, X, rows, k, Ct, ldc, nc, out_pos, labels, best, few_rows ? cap + 1 : 0L,
                     LONG_MAX);
  return hipGetLastError();
}

hipError_t launch_assign_f64(hipStream_t s, long m, int d, const double *X, int k,
                             const double *Ct, int ldc, const double *nc, int *labels) {
  long tiles = km_cdiv(m, am_bm<double>());
  if (tiles <= 0) return hipSuccess;
  long grid = tiles < (long)num_cus() * 8 ? tiles : (long)num_cus() * 8;
  hipLaunchKernelGGL(k_assign_f64_mfma, dim3(grid), dim3(256), 0, s, m, (const int *)nullptr, d, X,
                     (const int *)nullptr, k, Ct, ldc, nc, (const int *)nullptr, labels,
                     (double *)nullptr);
  return hipGetLastError();
}

// Waves per scatter sub-block (a power of two dividing KM_SB): its W x K LDS
// counters stay within 64 KB.
static int scatter_waves(int k) { return k <= 1024 ? 16 : k <= 2048 ? 8 : 1; }
// Points per sub-block (one histogram column).
int sub_block(int) { return KM_SB; }


size_t update_ws_layout(long n, int d, int k, char *base, UpdateWs *ws) {
  const long nb = n < KM_BLOCK ? n : KM_BLOCK;
  const int SB = sub_block(k);
  const long nsb = km_cdiv(nb > 0 ? nb : 1, SB);
  const long L = (long)k * nsb;
  const long T = km_cdiv(L + 1, SCAN_TILE);
  const long qmax = nb / KM_CHUNK + k + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return base ? (void *)(base + o) : nullptr;
  };
  UpdateWs w;
  w.hist = (int *)take(sizeof(int) * L);
  w.offs = (int *)take(sizeof(int) * (L + 1));
  w.tsum = (int *)take(sizeof(int) * (T + 1));
  w.cstart = (int *)take(sizeof(int) * k);
  w.ccount = (int *)take(sizeof(int) * k);
  w.cq = (int *)take(sizeof(int) * (k + 1));
  w.perm = (int *)take(sizeof(int) * (nb > 0 ? nb : 1));
  w.dummy = (int *)take(sizeof(int) * SCS_THREADS);
  w.P = (double *)take(sizeof(double) * qmax * (d + 1));
  w.S = (double *)take(sizeof(double) * (long)k * (d + 1));
  w.cnt = (long long *)take(sizeof(long long) * k);
  w.compj = (double *)take(sizeof(double) * (k + 1));
  w.bytes = off;
  w.perm_out = nullptr;
  if (ws) *ws = w;
  return off;
}

static hipError_t scan_excl(hipStream_t s, const int *in, long L, int *tsum, int *out) {
  const long T = km_cdiv(L + 1, SCAN_TILE);
  hipLaunchKernelGGL(k_scan_reduce, dim3(T), dim3(256), 0, s, in, L, tsum);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, tsum, (int)T);
  hipLaunchKernelGGL(k_scan_down, dim3(T), dim3(256), 0, s, in, L, tsum, out);
  return hipGetLastError();
}

long scan_tiles(long L) { return km_cdiv(L + 1, SCAN_TILE); }
hipError_t launch_scan_excl(hipStream_t s, const int *in, long L, int *tsum, int *out) {
  return scan_excl(s, in, L, tsum, out);
}

hipError_t launch_point_sq(hipStream_t s, long m, int d, const float *X, double *q) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_point_sq, dim3(km_cdiv(m, 256)), dim3(256), 0, s, m, d, X, q);
  return hipGetLastError();
}

hipError_t launch_point_sq(hipStream_t s, long m, int d, const double *X, double *q) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_point_sq_f64, dim3(km_cdiv(m, 256)), dim3(256), 0, s, m, d, X, q);
  return hipGetLastError();
}

hipError_t launch_update_zero(hipStream_t s, int d, int k, double *S, long long *cnt, int *empty) {
  const long zn = (long)k * (d + 1);
  const long zmax = zn > k ? zn : k;
  hipLaunchKernelGGL(k_zero_sums, dim3(km_cdiv(zmax, 256)), dim3(256), 0, s, zn, S, k, cnt, empty);
  return hipGetLastError();
}

// The stable counting sort of one canonical block's labels (rows
// [i0, i0 + nb)): sub-block histograms, their exclusive scan, the scatter
// into perm (or ws.perm_out at the block's rows) and the per-centre ranges.
static hipError_t block_sort(hipStream_t s, long i0, long nb, int k, const int *labels,
                             const UpdateWs &ws) {
  hipError_t e;
  const int SB = sub_block(k), W = scatter_waves(k);
  int nbits = 0;
  while ((1 << nbits) < k) ++nbits;
  const int nsb = (int)km_cdiv(nb, SB);
  const long L = (long)k * nsb;
  hipLaunchKernelGGL(k_hist, dim3(nsb), dim3(SB >= 32768 ? 1024 : 256), sizeof(int) * k, s,
                     labels + i0, nb, k, nsb, SB, ws.hist);
  if ((e = scan_excl(s, ws.hist, L, ws.tsum, ws.offs)) != hipSuccess) return e;
  int *perm = ws.perm_out ? ws.perm_out + i0 : ws.perm;
  // the staged scatter while its LDS fits (k <= SCS_MAXK: 61 KB at 1024)
  // and the block has enough sub-blocks for its 8-wave workgroups to fill
  // the chip (measured: 98 sub-blocks 54 -> 29 us, 7 sub-blocks 13.6 ->
  // 16.9 us); else the peer match when the counters and the masks fit 64 KB
  // of LDS with the same waves (fewer waves per sub-block cost more than the
  // ballots the masks save)
  const size_t lds_peer = sizeof(int) * ((size_t)W * k + 1) + sizeof(unsigned long long) * 64 * W;
  if (k <= SCS_MAXK && nsb >= SCS_MIN_SUBBLOCKS) {
    const size_t lds = sizeof(int) * (SCS_TILE + (size_t)((k + 1) & ~1) + ((k + 2) & ~1)) +
                       sizeof(unsigned long long) * (SCS_THREADS / 64) * 64 +
                       sizeof(uint16_t) * (SCS_THREADS / 64) * (size_t)k;
    hipLaunchKernelGGL(k_scatter_staged, dim3(nsb), dim3(SCS_THREADS), lds, s, labels + i0, i0, nb,
                       k, nsb, SB, nbits, ws.offs, perm, ws.dummy);
  } else if (lds_peer <= 65536)
    hipLaunchKernelGGL(k_scatter<true>, dim3(nsb), dim3(64 * W), lds_peer, s, labels + i0, i0, nb,
                       k, nsb, SB, SB / W, nbits, ws.offs, perm);
  else
    hipLaunchKernelGGL(k_scatter<false>, dim3(nsb), dim3(64 * W), sizeof(int) * W * k, s,
                       labels + i0, i0, nb, k, nsb, SB, SB / W, nbits, ws.offs, perm);
  hipLaunchKernelGGL(k_cluster_setup, dim3(1), dim3(1024), 0, s, k, nsb, ws.offs, ws.cstart,
                     ws.ccount, ws.cq);
  return hipGetLastError();
}

// The FP64 chunk sums of one sorted canonical block into ws.P.
template <typename T>
static hipError_t block_chunks(hipStream_t s, long i0, long nb, int d, int k, const T *X,
                               const UpdateWs &ws) {
  const int *perm = ws.perm_out ? ws.perm_out + i0 : ws.perm;
  const long qmax = nb / KM_CHUNK + k + 1;
  const bool reuse = ws.fd && ws.perm_prev;
  if (reuse)  // each centre's first member position that differs from the previous sort
    hipLaunchKernelGGL(k_first_diff, dim3(k), dim3(256), 0, s, k, perm, ws.cstart, ws.ccount,
                       ws.perm_prev + i0, ws.cs_prev, ws.fd);
  const int *fd = reuse ? ws.fd : nullptr;
  if (sizeof(T) == 4 && d % 4 == 0 && d >= 16 && d <= 256 && ((uintptr_t)X & 15) == 0) {
    int TPC = 1;
    while (TPC < d / 4) TPC <<= 1;
    const int G4 = 256 / TPC < 64 ? 256 / TPC : 64;  // must match k_chunk_sums4
    hipLaunchKernelGGL(k_chunk_sums4, dim3(km_cdiv(qmax, G4)), dim3(256), 0, s,
                       reinterpret_cast<const float *>(X), d, TPC, perm, k, ws.cstart, ws.ccount,
                       ws.cq, ws.P, fd, ws.cs_prev, ws.P_prev);
  } else {
    int Dp = 1;
    while (Dp < d && Dp < 256) Dp <<= 1;
    const int G = 256 / Dp < 64 ? 256 / Dp : 64;  // must match k_chunk_sums
    hipLaunchKernelGGL(k_chunk_sums<T>, dim3(km_cdiv(qmax, G)), dim3(256), 0, s, X, d, Dp, perm,
                       k, ws.cstart, ws.ccount, ws.cq, ws.P, fd, ws.cs_prev, ws.P_prev);
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t block_sums(hipStream_t s, long i0, long nb, int d, int k, const T *X,
                             const int *labels, const UpdateWs &ws) {
  hipError_t e;
  if ((e = block_sort(s, i0, nb, k, labels, ws)) != hipSuccess) return e;
  return block_chunks(s, i0, nb, d, k, X, ws);
}

// S[j] += the block's CSUM per centre (the order-dependent step: blocks are
// added left to right), cnt[j] += its member count.
static hipError_t block_add(hipStream_t s, int d, int k, double *S, long long *cnt,
                            const UpdateWs &ws) {
  hipLaunchKernelGGL(k_block_finalize, dim3(k), dim3(64 * (int)km_cdiv(d + 1 < 256 ? d + 1 : 256, 64)),
                     0, s, d, k, ws.ccount, ws.cq, ws.P, S, cnt);
  return hipGetLastError();
}

// Adds the canonical block sums of rows [0, n) of X (labels aligned with X)
// into S (k x (d+1) FP64: d sums and the squared-norm column) and cnt;
// blocks start at row 0 of X.
template <typename T>
static hipError_t launch_update_accumulate_t(hipStream_t s, long n, int d, int k, const T *X,
                                    const int *labels, double *S, long long *cnt,
                                    const UpdateWs &ws) {
  hipError_t e;
  for (long i0 = 0; i0 < n; i0 += KM_BLOCK) {
    const long nb = (n - i0) < KM_BLOCK ? (n - i0) : KM_BLOCK;
    if ((e = block_sums(s, i0, nb, d, k, X, labels, ws)) != hipSuccess) return e;
    if ((e = block_add(s, d, k, S, cnt, ws)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// The same sums with the blocks' sorts and chunk sums on two side streams
// (block b on side[b & 1] with workspace ws[b & 1]) and the additions into
// S on s in block order; the result is the single-stream one bit for bit.
// Blocks go in pairs: the two latency-bound sorts run side by side, and
// neither pair member's bandwidth-bound chunk sums start before both sorts
// are done, so they do not slow the sorts down.  A side workspace is reused
// two blocks later, after s has added the earlier block.
hipError_t launch_update_accumulate_2s(hipStream_t s, const UpdateOverlap &ov, long n, int d,
                                       int k, const float *X, const int *labels, double *S,
                                       long long *cnt, const UpdateWs *blk) {
  hipError_t e;
  if ((e = hipEventRecord(ov.start, s)) != hipSuccess) return e;
  for (int i = 0; i < 2; ++i)
    if ((e = hipStreamWaitEvent(ov.side[i], ov.start, 0)) != hipSuccess) return e;
  const long nblk = km_cdiv(n, KM_BLOCK);
  for (long p = 0; p < nblk; p += 2) {
    const bool pair = p + 1 < nblk;
    for (long b = p; b < p + 2 && b < nblk; ++b) {
      const int i = (int)(b & 1);
      const long i0 = b * KM_BLOCK, nb = std::min<long>(KM_BLOCK, n - i0);
      if (b >= 2 && (e = hipStreamWaitEvent(ov.side[i], ov.fin[i], 0)) != hipSuccess) return e;
      if ((e = block_sort(ov.side[i], i0, nb, k, labels, blk ? blk[b] : *ov.ws[i])) != hipSuccess)
        return e;
      if ((e = hipEventRecord(ov.sorted[i], ov.side[i])) != hipSuccess) return e;
    }
    for (long b = p; b < p + 2 && b < nblk; ++b) {
      const int i = (int)(b & 1);
      const long i0 = b * KM_BLOCK, nb = std::min<long>(KM_BLOCK, n - i0);
      if (pair && (e = hipStreamWaitEvent(ov.side[i], ov.sorted[i ^ 1], 0)) != hipSuccess)
        return e;
      if ((e = block_chunks(ov.side[i], i0, nb, d, k, X, blk ? blk[b] : *ov.ws[i])) != hipSuccess)
        return e;
      if ((e = hipEventRecord(ov.done[i], ov.side[i])) != hipSuccess) return e;
    }
    for (long b = p; b < p + 2 && b < nblk; ++b) {
      const int i = (int)(b & 1);
      if ((e = hipStreamWaitEvent(s, ov.done[i], 0)) != hipSuccess) return e;
      if ((e = block_add(s, d, k, S, cnt, blk ? blk[b] : *ov.ws[i])) != hipSuccess) return e;
      if ((e = hipEventRecord(ov.fin[i], s)) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

template <typename T>
static hipError_t launch_update_finalize_t(hipStream_t s, int d, int k, const double *S,
                                  const long long *cnt, const T *C_old, T *C_new, T *Ct_new,
                                  int ldc, T *nc_new, long long *counts, int *empty,
                                  double *compj) {
  const bool stage = compj && (sizeof(T) + sizeof(double)) * (size_t)d <= 65536;
  const size_t lds = sizeof(T) * d + (stage ? sizeof(double) * d : 0);
  hipLaunchKernelGGL(k_final_centres<T>, dim3(k), dim3(d < 256 ? 64 * (int)km_cdiv(d, 64) : 256),
                     lds, s, d, k, S, cnt, C_old, C_new, Ct_new, ldc, nc_new, counts, empty,
                     compj, (int)stage);
  return hipGetLastError();
}

// Full update; ws.compj receives the per-centre compactness terms against
// the new centres (reduce with launch_comp_reduce).
template <typename T>
static hipError_t launch_update_t(hipStream_t s, long n, int d, int k, const T *X, const int *labels,
                         const T *C_old, T *C_new, T *Ct_new, int ldc,
                         T *nc_new, long long *counts, int *empty, const UpdateWs &ws) {
  hipError_t e;
  if ((e = launch_update_zero(s, d, k, ws.S, ws.cnt, empty)) != hipSuccess) return e;
  if ((e = launch_update_accumulate(s, n, d, k, X, labels, ws.S, ws.cnt, ws)) != hipSuccess)
    return e;
  return launch_update_finalize(s, d, k, ws.S, ws.cnt, C_old, C_new, Ct_new, ldc, nc_new, counts,
                                empty, ws.compj);
}

hipError_t launch_comp_reduce(hipStream_t s, int k, const double *compj, double *out) {
  hipLaunchKernelGGL(k_comp_reduce, dim3(1), dim3(256), 0, s, compj, k, out);
  return hipGetLastError();
}

// Compactness of labels against centres C (any centres): moments, then terms.
template <typename T>
static hipError_t launch_compactness_t(hipStream_t s, long n, int d, int k, const T *X,
                              const int *labels, const T *C, double *out, const UpdateWs &ws) {
  hipError_t e;
  if ((e = launch_update_zero(s, d, k, ws.S, ws.cnt, nullptr)) != hipSuccess) return e;
  if ((e = launch_update_accumulate(s, n, d, k, X, labels, ws.S, ws.cnt, ws)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_comp_terms<T>, dim3(km_cdiv(k, 256)), dim3(256), 0, s, d, k, ws.S, ws.cnt, C,
                     ws.compj);
  return launch_comp_reduce(s, k, ws.compj, out);
}

// nc[j] = fma chain over d of Cdk[d*ld + j]^2 (D x K layout, Normalize.cu:33-47).
template <typename T>
__global__ void k_col_norms_dk(int k, int d, int ld, const T *__restrict__ C, T *__restrict__ nc) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  T s = 0;
  for (int q = 0; q < d; ++q) {
    const T v = C[(long)q * ld + j];
    s = fma_t(v, v, s);
  }
  nc[j] = s;
}

template <typename T>
static hipError_t launch_col_norms_dk_t(hipStream_t s, int k, int d, int ld, const T *C, T *nc) {
  hipLaunchKernelGGL(k_col_norms_dk<T>, dim3(km_cdiv(k, 256)), dim3(256), 0, s, k, d, ld, C, nc);
  return hipGetLastError();
}

// rowTransformMinimum (RowTransformMinimum.cu:3-58): for each row i of the
// m x n product P (row-major; n = number of centres), the lowest j among the
// minima of v_ij = (T)((double)(rn[i] + cn[j]) - 2.0 * (double)P[i][j]); the
// first sum is taken in T and the rest in double, then rounded back to T.
// Only values below FLT_MAX count (the reference's start value for both
// types, :17); a row with none gets `none` (-1 in the reference).
// rn is read at the row's local index i and out at m0 + i, as in the
// reference kernel (:20, :57).  One wave per row, lanes stride over j (16
// bytes at a time when the rows are aligned) and the wave merges with the
// tie rule.
__device__ inline void load16(const float *p, float *v) {
  const float4 a = *reinterpret_cast<const float4 *>(p);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ inline void load16(const double *p, double *v) {
  const double2 a = *reinterpret_cast<const double2 *>(p);
  v[0] = a.x; v[1] = a.y;
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void k_row_transform_min(long m0, long m, int n,
                                                           const T *__restrict__ rn,
                                                           const T *__restrict__ cn,
                                                           const T *__restrict__ P,
                                                           int *__restrict__ out, int none) {
  const T big = (T)FLT_MAX;
  const int lane = threadIdx.x & 63;
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  for (long i = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += nw) {
    const T r = rn[i];
    T bv = big;
    int bi = -1;
    if (VEC) {  // 16-byte loads: a lane owns runs of 16/sizeof(T) consecutive j
      constexpr int W = 16 / sizeof(T);
      for (int j0 = lane * W; j0 < n; j0 += 64 * W) {
        T pv[W];
        load16(P + i * n + j0, pv);
#pragma unroll
        for (int u = 0; u < W; ++u) {
          const T v = (T)((double)(r + cn[j0 + u]) - 2.0 * (double)pv[u]);
          if (v < bv) { bv = v; bi = j0 + u; }
        }
      }
    } else {
      for (int j = lane; j < n; j += 64) {
        const T v = (T)((double)(r + cn[j]) - 2.0 * (double)P[i * n + j]);
        if (v < bv) { bv = v; bi = j; }
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const T ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov < bv || (ov == bv && (unsigned)oi < (unsigned)bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) out[m0 + i] = bi < 0 ? none : bi;
  }
}

template <typename T>
static hipError_t launch_row_transform_min_t(hipStream_t s, long m0, long m, int n, const T *rn,
                                    const T *cn, const T *P, int *out, int none) {
  if (m <= 0) return hipSuccess;
  // vector loads when every row starts 16-byte aligned
  if (n % (16 / sizeof(T)) == 0 && ((uintptr_t)P & 15) == 0)
    hipLaunchKernelGGL((k_row_transform_min<T, true>), dim3(km_grid(m * 64, 256)), dim3(256), 0, s,
                       m0, m, n, rn, cn, P, out, none);
  else
    hipLaunchKernelGGL((k_row_transform_min<T, false>), dim3(km_grid(m * 64, 256)), dim3(256), 0, s,
                       m0, m, n, rn, cn, P, out, none);
  return hipGetLastError();
}

hipError_t launch_generate(hipStream_t s, uint64_t seed, long i0, long n, int d, int ncomp,
                           const float *means, const float *sigmas, const double *cdf, float *X,
                           int *z) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_generate, dim3(km_cdiv(n, 256)), dim3(256), 0, s, seed, i0, n, d, ncomp,
                     means, sigmas, cdf, X, z);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_gather_rows_t(hipStream_t s, long n, int d, const T *X, const int *idx, T *out) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows<T>, dim3(km_grid(n * d, 256)), dim3(256), 0, s, n, d, X, idx,
                     out);
  return hipGetLastError();
}

// The float and double entry points of every templated launcher.
#define KM_OVERLOADS(T)                                                                          \
  hipError_t launch_col_norms(hipStream_t s, int k, int d, const T *C, T *nc) {                  \
    return launch_col_norms_t<T>(s, k, d, C, nc);                                                \
  }                                                                                              \
  hipError_t launch_row_norms(hipStream_t s, long m, int d, const T *X, T *out) {                \
    return launch_row_norms_t<T>(s, m, d, X, out);                                               \
  }                                                                                              \
  hipError_t launch_transpose(hipStream_t s, int k, int d, const T *C, T *Ct, int ldc) {         \
    return launch_transpose_t<T>(s, k, d, C, Ct, ldc);                                           \
  }                                                                                              \
  hipError_t launch_update_accumulate(hipStream_t s, long n, int d, int k, const T *X,           \
                                      const int *labels, double *S, long long *cnt,              \
                                      const UpdateWs &ws) {                                      \
    return launch_update_accumulate_t<T>(s, n, d, k, X, labels, S, cnt, ws);                     \
  }                                                                                              \
  hipError_t launch_update_finalize(hipStream_t s, int d, int k, const double *S,                \
                                    const long long *cnt, const T *C_old, T *C_new, T *Ct_new,   \
                                    int ldc, T *nc_new, long long *counts, int *empty,           \
                                    double *compj) {                                             \
    return launch_update_finalize_t<T>(s, d, k, S, cnt, C_old, C_new, Ct_new, ldc, nc_new,       \
                                       counts, empty, compj);                                    \
  }                                                                                              \
  hipError_t launch_update(hipStream_t s, long n, int d, int k, const T *X, const int *labels,   \
                           const T *C_old, T *C_new, T *Ct_new, int ldc, T *nc_new,              \
                           long long *counts, int *empty, const UpdateWs &ws) {                  \
    return launch_update_t<T>(s, n, d, k, X, labels, C_old, C_new, Ct_new, ldc, nc_new,          \
                              counts, empty, ws);                                                \
  }                                                                                              \
  hipError_t launch_compactness(hipStream_t s, long n, int d, int k, const T *X,                 \
                                const int *labels, const T *C, double *out,                      \
                                const UpdateWs &ws) {                                            \
    return launch_compactness_t<T>(s, n, d, k, X, labels, C, out, ws);                           \
  }                                                                                              \
  hipError_t launch_col_norms_dk(hipStream_t s, int k, int d, int ld, const T *C, T *nc) {       \
    return launch_col_norms_dk_t<T>(s, k, d, ld, C, nc);                                         \
  }                                                                                              \
  hipError_t launch_row_transform_min(hipStream_t s, long m0, long m, int n, const T *rn,        \
                                      const T *cn, const T *P, int *out, int none) {             \
    return launch_row_transform_min_t<T>(s, m0, m, n, rn, cn, P, out, none);                     \
  }                                                                                              \
  hipError_t launch_gather_rows(hipStream_t s, long n, int d, const T *X, const int *idx,        \
                                T *out) {                                                        \
    return launch_gather_rows_t<T>(s, n, d, X, idx, out);                                        \
  }
KM_OVERLOADS(float)
KM_OVERLOADS(double)
#undef KM_OVERLOADS

}  // namespace km
