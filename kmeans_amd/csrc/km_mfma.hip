// Reduced-precision assignment on MFMA with an FP32 certificate.
//
// Labels are those of the plain FP32 path (main/cuda/ClosestCenters.cu:46-123
// semantics, fma-chain scores, lowest index among maxima).  The MFMA pass
// computes approximate scores t_ij = xh_i . ch_j - 0.5*nc_j (BF16 or FP16
// operands, FP32 accumulation) for every centre and keeps, per point, the
// three largest approximate scores and the indices of the first two.  The
// epilogue recomputes the exact FP32 chain for those two candidates; the
// point is certified when the better exact score exceeds the third
// approximate score plus a forward-error bound E_i (below), because then no
// other centre can reach it.  Uncertified points go to the FP32 kernel.
//
// Error bound (Higham, Accuracy and Stability of Numerical Algorithms, 2nd
// ed., sec. 3.1; u = 2^-24, gamma_n = n u / (1 - n u)), in FP32 score units
// (the FP16 scale s divided out), with dx = x - xh/s and dc = c - ch/s the
// representation errors of the reduced-precision operands:
//   |fl-chain - exact| <= gamma_D (0.5 nc_j + |x| |c_j|)          (FP32 path)
//   |xh.ch/s^2 - x.c|  =  |dx.c + (xh/s).dc|
//                      <= |dx| |c_j| + (|x| + |dx|) |dc_j|        (rounding)
//   |mfma - xh.ch|/s^2 <= gamma'_{D+1} (0.5 nc_j + (|x| + |dx|)(|c_j| + |dc_j|))
//                         with a pessimistic unit roundoff 2^-22 for the
//                         matrix core
// (Cauchy-Schwarz), maximised over j with the largest |c_j|, nc_j and
// |dc_j|.  |dx| is measured per point and |dc_j| per centre (norms of the
// actual rounding errors, typically 0.3-0.4 ub |x| for ub = 2^-8 (BF16) or
// 2^-11 (FP16), against the worst case ub |x| of a component-wise bound).
#include <float.h>
#include <math.h>

#include <vector>

#include "km_engine.hpp"
#include "km_mfma.hpp"

namespace km {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4_t;

#define MF_TILE 32     // centres per MFMA tile (rows of the 32x32 output)
#ifndef MF_WAVES
#define MF_WAVES 4     // waves per workgroup; every wave reads the same centre stages
#endif
#define MF_PTS (64 * MF_WAVES)  // points per workgroup (2 groups of 32 per wave)
#define MF_STAGE 2     // centre tiles per LDS stage (64 centres): 33 KB of LDS a
                       // workgroup, so four fit on a CU
#define MF_KPAD 128    // centres are padded to a multiple of this
// rescored-point counter of one assign launch: 64 counters on separate
// 128-byte lines (workgroup b adds into counter b % 64), summed on the host
#define NC_SPREAD 64
#define NC_STRIDE 32
#define NC_SLOT (NC_SPREAD * NC_STRIDE)
// largest k with the label-ordered assign (its tile table costs k^2 d flops
// per centre update)
#define MF_ORD_MAX_K 8192

// Round-to-nearest-even conversions (the hardware instructions; the
// certificate only needs the rounded values to be those the MFMAs see).
__device__ inline uint16_t f32_to_bf16_rne(float f) { return bf16_rne(f); }
__device__ inline uint16_t f32_to_f16_rne(float f) {
  _Float16 h = (_Float16)f;  // IEEE conversion, round to nearest even
  return __builtin_bit_cast(uint16_t, h);
}

struct MfmaState {
  int Dp = 0;        // padded dimension (multiple of 16)
  int KS = 0;        // k-steps of 16
  int Kp = 0;        // centres padded to MF_KPAD
  bool fp16 = false;
  float scale = 1.f; // power of two applied to X and C before conversion (FP16)
  uint16_t *Xh = nullptr;   // m x Dp
  uint16_t *Ch = nullptr;   // Kp x Dp
  uint16_t *Cl = nullptr;   // Kp x Dp: BF16 of (C - Ch), the low part (BF16 engines)
  // the deep pass's unscaled BF16 centre parts and biases: Ch, Cl and bias
  // themselves for BF16 engines, own copies for FP16 engines
  uint16_t *Chb = nullptr, *Clb = nullptr;
  float *biasb = nullptr;
  float *bias = nullptr;    // Kp: -0.5*scale^2*nc (or -FLT_MAX past k)
  float *bnd = nullptr;     // [cmax, ncmax, dcmax]: bounds over the centres
  float *dcn = nullptr;     // Kp: upper bound of |c_j - ch_j/s|
  float2 *xnb = nullptr;    // m: upper bounds of |x_i| and |x_i - xh_i/s|
  int *flags = nullptr;     // m
  int *ncand = nullptr;     // per iteration slot: NC_SPREAD counters, NC_STRIDE ints apart
  int *nflag = nullptr;     // per iteration slot
  int *flags2 = nullptr;    // m: points left open by the deep pass
  int *nflag2 = nullptr;    // per iteration slot
  unsigned long long *kbest = nullptr;  // split_rows_cap() keys of the FP32 search, zeroed
  // label-ordered assign: ttab[j][0..nKT) = the centre tiles by increasing
  // distance from centre j (k <= MF_ORD_MAX_K), home[wg] = the previous label
  // of a workgroup's first point
  int *ttab = nullptr;
  float *tdist = nullptr;   // tdist[j][rank]: lower bound of |c_j - c_k| over tile ttab[j][rank]
  int *home = nullptr;
  long home_cap = 0;
  bool ttab_ok = false;
  // drift bounds of the resident points: xb[i] = (u, l, l3, r) with u >=
  // |x_i - c_label| and l <= |x_i - c_j| for every other centre j, r (int
  // bits, -1 for none) the runner-up centre and l3 <= |x_i - c_j| for every
  // centre j other than the label and r, valid for the centres Cref;
  // drift[j] >= |c_j - Cref_j| and drift[k] their maximum.  ulist: the
  // points the bounds leave open (from the front) and the two-candidate
  // points (from the back), fcnt their counts.  bnd_valid: xb and Cref
  // describe the last assign.
  float4 *xb = nullptr;
  float *Cref = nullptr, *drift = nullptr;
  int *ulist = nullptr, *fcnt = nullptr;
  bool bnd_valid = false;
  // the two-candidate decisions run on a side stream (default priority)
  // after the assign kernel, beside the re-resolve: they touch only points
  // the assign and the re-resolve skip; pair_pending until the engine's
  // stream has waited for them (mfma_pair_join)
  hipStream_t pair_stream = nullptr;
  hipEvent_t pair_fork = nullptr, pair_join = nullptr;
  bool pair_pending = false;
  int nflag_slots = 0;
  bool x_ready = false;
  bool c_ready = false;
  // timing runs: start/stop events of each assign launch not yet collected
  std::vector<hipEvent_t> kev;
  size_t kev_used = 0;
};

static MfmaState *st(Engine *e) { return (MfmaState *)e->mfma; }
static void collect_kernel_times(Engine *e, MfmaState *s);
static hipError_t timing_begin(Engine *e, MfmaState *s, hipEvent_t *k1);
static hipError_t timing_end(Engine *e, hipEvent_t k1);
static hipError_t launch_assign(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                                int *lab, long rows, const int *order, const int *home,
                                const int *Mdev, float4 *xbo);

// ------------------------------------------------------------------ kernels
__global__ void k_convert_rows(long m, int d, int Dp, const float *__restrict__ X, float scale,
                               int fp16, uint16_t *__restrict__ out) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < m * Dp;
       e += (long)gridDim.x * blockDim.x) {
    const long r = e / Dp;
    const int q = (int)(e - r * Dp);
    const float v = q < d ? X[r * d + q] * scale : 0.0f;
    out[e] = fp16 ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
  }
}

// Ch (and for BF16 the low parts Cl), bias and the per-centre bound dcn_j
// >= |c_j - ch_j/s|: one wave per centre row.  c*s - float(ch) is exact in
// FP32; the FP32 sum of its squares is within gamma_Dp of the exact sum.
__global__ __launch_bounds__(256) void k_prepare_centres(int k, int d, int Dp, int Kp,
                                                         const float *__restrict__ C,
                                                         const float *__restrict__ nc, float scale,
                                                         int fp16, uint16_t *__restrict__ Ch,
                                                         uint16_t *__restrict__ Cl,
                                                         float *__restrict__ bias,
                                                         float *__restrict__ dcn,
                                                         uint16_t *__restrict__ Chb,
                                                         uint16_t *__restrict__ Clb,
                                                         float *__restrict__ biasb) {
  const int lane = threadIdx.x & 63;
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  for (long j = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < Kp; j += nw) {
    float sq = 0.0f;
    for (int q = lane; q < Dp; q += 64) {
      const float v = (j < k && q < d) ? C[j * d + q] * scale : 0.0f;
      const uint16_t hi = fp16 ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
      const float hv = fp16 ? (float)__builtin_bit_cast(_Float16, hi) : __uint_as_float((uint32_t)hi << 16);
      const float r = v - hv;
      Ch[j * Dp + q] = hi;
      if (!fp16) Cl[j * Dp + q] = f32_to_bf16_rne(r);
      if (fp16) {  // unscaled BF16 parts for the deep pass
        const float c = (j < k && q < d) ? C[j * d + q] : 0.0f;
        const uint16_t bh = f32_to_bf16_rne(c);
        Chb[j * Dp + q] = bh;
        Clb[j * Dp + q] = f32_to_bf16_rne(c - __uint_as_float((uint32_t)bh << 16));
      }
      sq = __builtin_fmaf(r, r, sq);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
    if (lane == 0) {
      bias[j] = j < k ? -0.5f * nc[j] * (scale * scale) : -FLT_MAX;
      if (fp16) biasb[j] = j < k ? -0.5f * nc[j] : -FLT_MAX;
      dcn[j] = sqrtf(sq * (1.0f + (float)Dp * 1.2e-7f + 1e-6f)) * (1.0f + 1e-6f) / scale + 1e-30f;
    }
  }
}

// bnd = [cmax, ncmax, dcmax]: upper bounds of max |c_j|, max |c_j|^2 and
// max |c_j - ch_j/s| over the k centres (one block).  nc_j is the FP32 chain
// of squares; the true |c_j|^2 is at most nc_j (1 + gamma_d).
__global__ __launch_bounds__(256) void k_centre_bounds(int k, int d, const float *__restrict__ nc,
                                                       const float *__restrict__ dcn,
                                                       float *__restrict__ bnd) {
  __shared__ float r1[256], r2[256];
  float m1 = 0.0f, m2 = 0.0f;
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    m1 = fmaxf(m1, nc[j]);
    m2 = fmaxf(m2, dcn[j]);
  }
  r1[threadIdx.x] = m1;
  r2[threadIdx.x] = m2;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      r1[threadIdx.x] = fmaxf(r1[threadIdx.x], r1[threadIdx.x + off]);
      r2[threadIdx.x] = fmaxf(r2[threadIdx.x], r2[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float ncmax = r1[0] * (1.0f + (float)d * 1.2e-7f + 1e-6f) + 1e-30f;
    bnd[0] = sqrtf(ncmax) * (1.0f + 1e-6f);
    bnd[1] = ncmax;
    bnd[2] = r2[0];
  }
}

// ttab[j][rank] = the centre tiles (MF_TILE rows each) in increasing order of
// a lower bound of the smallest squared distance from centre j to a member
// (ties to the lower tile), one workgroup per centre, and tdist[j][rank] =
// the square root of that bound (+inf for a tile of padding rows).  The
// squared distance is nc_j + nc_i - 2 c_j.c_i in FP32 (fma chain for the dot
// product); nc are FP32 chains of squares.  Each of the three is within
// gamma_D of its exact value, relative to nc_j + nc_i (|c_j.c_i| is at most
// half of it), and the two subtractions add at most 2 ulps of that size, so
// subtracting (2 D + 8) 2^-24 (nc_j + nc_i) (with a 1% margin) gives a bound.
__global__ __launch_bounds__(256) void k_tile_order(int K, int D, int nKT,
                                                    const float *__restrict__ Ct, int ldc,
                                                    const float *__restrict__ nc,
                                                    int *__restrict__ ttab,
                                                    float *__restrict__ tdist) {
  extern __shared__ float to_smem[];
  float *cj = to_smem;                                              // [D]
  unsigned *key = reinterpret_cast<unsigned *>(to_smem + D);       // [nKT]
  const int j = blockIdx.x, tid = threadIdx.x;
  for (int d = tid; d < D; d += 256) cj[d] = Ct[(long)d * ldc + j];
  for (int t = tid; t < nKT; t += 256) key[t] = 0x7F800000u;  // +inf
  __syncthreads();
  const float ncj = nc[j];
  for (int i = tid; i < K; i += 256) {
    float dot = 0.0f;
    for (int d = 0; d < D; ++d) dot = __builtin_fmaf(cj[d], Ct[(long)d * ldc + i], dot);
    const float err = (2.0f * (float)D + 8.0f) * 5.9604645e-8f * 1.01f * (ncj + nc[i]);
    const float d2 = fmaxf((ncj + nc[i] - 2.0f * dot) - err, 0.0f);
    atomicMin(&key[i / MF_TILE], __float_as_uint(d2));  // non-negative: bits order as values
  }
  __syncthreads();
  for (int t = tid; t < nKT; t += 256) {
    const unsigned kt = key[t];
    int r = 0;
    for (int u = 0; u < nKT; ++u) r += (key[u] < kt) || (key[u] == kt && u < t);
    ttab[(long)j * nKT + r] = t;
    tdist[(long)j * nKT + r] = sqrtf(__uint_as_float(kt)) * (1.0f - 1e-6f);
  }
}

// ---------------------------------------------------------------- drift bounds
// A point keeps its label without being scored when its bounds, moved by the
// centres' drift since they were made, still separate the label from every
// other centre (Hamerly, "Making k-means even faster", SDM 2010, with the
// separation required to exceed the FP32 chains' error): with u' = u +
// drift[a] >= |x - c_a| and l' = l - max drift <= |x - c_j| (triangle
// inequality), the exact scores differ by 0.5 (|x - c_j|^2 - |x - c_a|^2) >=
// 0.5 (l'^2 - u'^2), and l'^2 - u'^2 > 4 E32 (E32 the FP32 chain bound of
// the single-problem certificate) makes the FP32 chain of a larger than
// every other, strictly, so the label is the plain path's.

// drift[j] = upper bound of |C_j - Cref_j| (the FP32 differences are within
// an ulp, the sum of squares within gamma_D in any order), one wave per
// centre; drift[K] (as bits) += the largest by atomic maximum over
// non-negative floats, from 0 (k_ref_copy clears it for the next pass);
// drift_bound: 0 exactly when every coordinate difference is 0.
__global__ __launch_bounds__(256) void k_drift(int K, int D, const float *__restrict__ C,
                                               const float *__restrict__ Cref,
                                               float *__restrict__ drift, int *__restrict__ fcnt,
                                               const int *__restrict__ stop) {
  if (blockIdx.x == 0 && threadIdx.x < 2) fcnt[threadIdx.x] = 0;  // the filter's counts
  if (stop && *stop) return;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= K) return;
  float q = 0.0f, ma = 0.0f;
  for (int d = lane; d < D; d += 64) {
    const float t = C[(long)j * D + d] - Cref[(long)j * D + d];
    q = __builtin_fmaf(t, t, q);
    const float at = fabsf(t);
    if (!(at <= ma)) ma = at;  // NaN propagates
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    q += __shfl_xor(q, o);
    const float om = __shfl_xor(ma, o);
    if (!(om <= ma)) ma = om;
  }
  if (lane == 0) {
    const float v = drift_bound(q, ma, D);
    drift[j] = v;
    atomicMax(reinterpret_cast<unsigned *>(drift + K), __float_as_uint(v));
  }
}

// The plain path's score chains of one point against two centres (-0.5 nc_j,
// then one fused multiply-add per dimension in ascending order), one pass
// over the point's row.
__device__ inline void chain2(const float *__restrict__ xr, const float *__restrict__ ra,
                              const float *__restrict__ rb, int D, bool vec, float &sa,
                              float &sb) {
  if (vec) {  // D % 4 == 0 and 16-byte aligned rows
    const float4 *x4 = reinterpret_cast<const float4 *>(xr);
    const float4 *a4 = reinterpret_cast<const float4 *>(ra);
    const float4 *b4 = reinterpret_cast<const float4 *>(rb);
#pragma unroll 4
    for (int q = 0; q < D / 4; ++q) {
      const float4 xv = x4[q], av = a4[q], bv = b4[q];
      sa = __builtin_fmaf(xv.x, av.x, sa);
      sb = __builtin_fmaf(xv.x, bv.x, sb);
      sa = __builtin_fmaf(xv.y, av.y, sa);
      sb = __builtin_fmaf(xv.y, bv.y, sb);
      sa = __builtin_fmaf(xv.z, av.z, sa);
      sb = __builtin_fmaf(xv.z, bv.z, sb);
      sa = __builtin_fmaf(xv.w, av.w, sa);
      sb = __builtin_fmaf(xv.w, bv.w, sb);
    }
  } else {
    for (int q = 0; q < D; ++q) {
      sa = __builtin_fmaf(xr[q], ra[q], sa);
      sb = __builtin_fmaf(xr[q], rb[q], sb);
    }
  }
}

// The test for a point with label a, bounds b = (u, l, l3, r), drift da of
// its label's centre and dr of r's, row norm bound xn and largest drift dm;
// its bounds moved by the drift (u' = u + da, l' = l - dm, l3' = l3 - dm):
// 0 when they keep its label (l'^2 - u'^2 > 4 E32 as above), 1 when
// they separate every centre but the label a and the runner-up r in the same
// way (l3'^2 - u'^2 > 4 E32: the FP32 chain of a then exceeds every other
// centre's but possibly r's, so the label is a or r, whichever chain is
// larger, lowest index on equal chains), 2 when the point must be scored.
// When a and r have not moved at all (drift 0: every coordinate equal, so
// both chains are the ones the last assign compared, which put a first) the
// second case keeps the label too.  nb = the moved bounds; the FP32 chains'
// error bound E32 comes from bnd = the centre bounds [cmax, ncmax, ...] of
// the current centres.
__device__ inline int bound_class(int a, float4 b, float da, float dr, float xn, float dm, int K,
                                  const float *bnd, float Df, float4 &nb) {
  const float u = (b.x + da) * (1.0f + 1e-6f);
  const float l = fmaxf((b.y - dm) * (1.0f - 1e-6f), 0.0f);
  const float l3 = fmaxf((b.z - dm) * (1.0f - 1e-6f), 0.0f);
  nb = make_float4(u, l, l3, b.w);
  const float e32 = Df * 5.9604645e-8f * 1.0001f * (0.5f * bnd[1] + xn * bnd[0]);
  const float t = 4.0f * e32 * 1.0001f + 1e-30f;
  if (l > u && (l - u) * (l + u) * (1.0f - 1e-6f) > t) return 0;
  const int r = __float_as_int(b.w);
  if (r >= 0 && r < K && r != a && l3 > u && (l3 - u) * (l3 + u) * (1.0f - 1e-6f) > t)
    return da == 0.0f && dr == 0.0f ? 0 : 1;
  return 2;
}

// One pass over the positions, BF_PPT per thread: the open points go to the
// front of ulist and the two-candidate points to its back (ulist[m - 1 - i]
// for the i-th), each workgroup reserving its runs with one atomic per list
// on fcnt = [open count, two-candidate count] (zeroed by k_drift); the moved
// bounds of every point not left open go to xb.  The lists' order among
// workgroups is the order of the reservations; every point is handled on
// its own, so no result depends on it.
#define BF_PPT 8
__global__ __launch_bounds__(256) void k_bound_filter(
    long M, const int *__restrict__ order, const int *__restrict__ labels, float4 *__restrict__ xb,
    const float *__restrict__ drift, int K, const float *__restrict__ bnd,
    const float2 *__restrict__ xnb, int D, long m, int *__restrict__ ulist,
    int *__restrict__ fcnt, const int *__restrict__ stop) {
  if (stop && *stop) return;
  __shared__ int wc[2][BF_PPT][4];
  __shared__ int base_s[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const long q0 = (long)blockIdx.x * 256 * BF_PPT;
  int cls[BF_PPT], pv[BF_PPT], la[BF_PPT];
  float4 bv[BF_PPT];
  float xv[BF_PPT], da[BF_PPT], dr[BF_PPT];
  uint64_t mo[BF_PPT], mp[BF_PPT];
  // every load of the thread's points first, then the drifts of their
  // centres, so each thread waits for memory twice
#pragma unroll
  for (int u = 0; u < BF_PPT; ++u) {
    const long q = q0 + u * 256 + threadIdx.x;
    const long p = q < M ? (order ? (long)order[q] : q) : 0;
    pv[u] = (int)p;
    la[u] = q < M ? labels[p] : 0;
    bv[u] = q < M ? xb[p] : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    xv[u] = q < M ? xnb[p].x : 0.0f;
  }
  const float dm = drift[K];
#pragma unroll
  for (int u = 0; u < BF_PPT; ++u) {
    const int r = __float_as_int(bv[u].w);
    da[u] = drift[la[u]];
    dr[u] = (r >= 0 && r < K) ? drift[r] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < BF_PPT; ++u) {
    const long q = q0 + u * 256 + threadIdx.x;
    cls[u] = 0;
    if (q < M) {
      float4 nb;
      cls[u] = bound_class(la[u], bv[u], da[u], dr[u], xv[u], dm, K, bnd, (float)D, nb);
      if (cls[u] != 2) xb[pv[u]] = nb;
    }
    mo[u] = __ballot(q < M && cls[u] == 2);
    mp[u] = __ballot(q < M && cls[u] == 1);
    if (lane == 0) {
      wc[0][u][w] = __popcll(mo[u]);
      wc[1][u][w] = __popcll(mp[u]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 2) {  // the workgroup's runs in the two lists
    int t = 0;
    for (int u = 0; u < BF_PPT; ++u)
      for (int v = 0; v < 4; ++v) t += wc[threadIdx.x][u][v];
    base_s[threadIdx.x] = t ? atomicAdd(fcnt + threadIdx.x, t) : 0;
  }
  __syncthreads();
  int bo = base_s[0], bp = base_s[1];
#pragma unroll
  for (int u = 0; u < BF_PPT; ++u) {
    int so = bo, sp = bp;  // this wave's offsets in round u
    for (int v = 0; v < w; ++v) {
      so += wc[0][u][v];
      sp += wc[1][u][v];
    }
    if (cls[u] == 2 && ((mo[u] >> lane) & 1)) ulist[so + __popcll(mo[u] & below)] = pv[u];
    if (cls[u] == 1 && ((mp[u] >> lane) & 1)) ulist[m - 1 - (sp + __popcll(mp[u] & below))] = pv[u];
    for (int v = 0; v < 4; ++v) {
      bo += wc[0][u][v];
      bp += wc[1][u][v];
    }
  }
}

// The two-candidate points (the last *pcnt of the m entries of ulist; xb
// holds their moved bounds (u', l', l3', r)): the label
// is the label a or the runner-up r, whichever FP32 chain is larger (lowest
// index on equal chains; X, C, nc the rows, the current centres and their
// chains of squares), since every other centre is separated by the bounds.
// Fresh bounds from the two chains (each within e32 of its exact score),
// l3' kept; the count goes to slot 4 of the counter lines ncl.
// Each wave takes 64 points, one per lane.  With 16-byte rows (D % 4 == 0,
// aligned) the wave stages its points' rows and both candidates' centre rows
// in LDS PR_PD dimensions at a time with coalesced loads (PR_PD / 4 lanes
// per piece of a row) and each lane continues its two chains from there;
// otherwise each lane reads its rows itself.
#define PR_PD 16                 // dimensions per staged piece
#define PR_PS (PR_PD / 4 + 1)    // float4 stride of a staged row (conflict-free reads)
__global__ __launch_bounds__(256) void k_pair_decide(
    const int *__restrict__ ulist, const int *__restrict__ pcnt, long m,
    int *__restrict__ labels, float4 *__restrict__ xb, const float2 *__restrict__ xnb,
    const float *__restrict__ bnd, int D, const float *__restrict__ X,
    const float *__restrict__ C, const float *__restrict__ nc, int *__restrict__ ncl,
    const int *__restrict__ stop) {
  if (stop && *stop) return;
  __shared__ float4 lds[4 * 3 * 64 * PR_PS];  // per wave: point rows, label rows, runner-up rows
  const long i0 = m - *pcnt, i1 = m;
  const float Df = (float)D;
  const bool vec = (D & 3) == 0 && (((uintptr_t)X | (uintptr_t)C) & 15) == 0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float4 *lx = lds + w * 3 * 64 * PR_PS, *la = lx + 64 * PR_PS, *lr = la + 64 * PR_PS;
  const long nw = (long)gridDim.x * 4;
  int cnt = 0;
  for (long g = (long)blockIdx.x * 4 + w; i0 + g * 64 < i1; g += nw) {
    const long i = i0 + g * 64 + lane;
    const bool ok = i < i1;
    const long p = ok ? ulist[i] : 0;
    const int a = ok ? labels[p] : 0;
    const float4 b = ok ? xb[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int r = ok ? __float_as_int(b.w) : 0;
    float sa = ok ? -0.5f * nc[a] : 0.0f, sr = ok ? -0.5f * nc[r] : 0.0f;
    if (vec) {
      constexpr int LPR = PR_PD / 4, RPI = 64 / LPR;  // lanes per row piece, rows per load
      const int np = (D + PR_PD - 1) / PR_PD;
      const int q = lane % LPR;
      // row t*RPI + lane/LPR of each array, 16-byte piece lane % LPR: all
      // loads of a piece in flight together, and the next piece's loads
      // issued before this piece's chains run
      float4 xv[LPR], av[LPR], bv[LPR];
      auto load_piece = [&](int pc) {
        const int d0 = pc * PR_PD + q * 4;
#pragma unroll
        for (int t = 0; t < LPR; ++t) {
          const int row = t * RPI + lane / LPR;
          const long pp = __shfl(p, row);
          const int aa = __shfl(a, row), rr = __shfl(r, row);
          const bool v = i0 + g * 64 + row < i1 && d0 < D;
          xv[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          av[t] = xv[t];
          bv[t] = xv[t];
          if (v) {
            xv[t] = *reinterpret_cast<const float4 *>(X + pp * D + d0);
            av[t] = *reinterpret_cast<const float4 *>(C + (long)aa * D + d0);
            bv[t] = *reinterpret_cast<const float4 *>(C + (long)rr * D + d0);
          }
        }
      };
      load_piece(0);
      for (int pc = 0; pc < np; ++pc) {
#pragma unroll
        for (int t = 0; t < LPR; ++t) {
          const int row = t * RPI + lane / LPR;
          lx[row * PR_PS + q] = xv[t];
          la[row * PR_PS + q] = av[t];
          lr[row * PR_PS + q] = bv[t];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (pc + 1 < np) load_piece(pc + 1);  // uniform
        // the chains continue over this piece's dimensions, in order
#pragma unroll
        for (int u = 0; u < PR_PD / 4; ++u) {
          if (pc * PR_PD + u * 4 < D) {  // uniform
            const float4 x4 = lx[lane * PR_PS + u], a4 = la[lane * PR_PS + u],
                         b4 = lr[lane * PR_PS + u];
            sa = __builtin_fmaf(x4.x, a4.x, sa);
            sr = __builtin_fmaf(x4.x, b4.x, sr);
            sa = __builtin_fmaf(x4.y, a4.y, sa);
            sr = __builtin_fmaf(x4.y, b4.y, sr);
            sa = __builtin_fmaf(x4.z, a4.z, sa);
            sr = __builtin_fmaf(x4.z, b4.z, sr);
            sa = __builtin_fmaf(x4.w, a4.w, sa);
            sr = __builtin_fmaf(x4.w, b4.w, sr);
          }
        }
        // the next piece overwrites the staging arrays
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    } else if (ok) {
      chain2(X + p * D, C + (long)a * D, C + (long)r * D, D, false, sa, sr);
    }
    if (ok) {
      const float xn = xnb[p].x;
      const float e32 = Df * 5.9604645e-8f * 1.0001f * (0.5f * bnd[1] + xn * bnd[0]);
      const bool tr = sr > sa || (sr == sa && r < a);
      const int L = tr ? r : a, O = tr ? a : r;
      const float sL = tr ? sr : sa, sO = tr ? sa : sr;
      labels[p] = L;
      const float lO = dist_lower(xn, sO + e32, Df);
      xb[p] = make_float4(dist_upper(xn, sL - e32), fminf(lO, b.z), b.z, __int_as_float(O));
      ++cnt;
    }
  }
  // per-wave sum, one atomic per wave with points
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0 && cnt) atomicAdd(ncl + (blockIdx.x & (NC_SPREAD - 1)) * NC_STRIDE + 4, cnt);
}

// Cref = C once the assign has made every point's bounds valid for C, and
// the drift maximum drift[K] back to 0 for the next k_drift
__global__ void k_ref_copy(long n, const float *__restrict__ C, float *__restrict__ Cref,
                           float *__restrict__ dmax, const int *__restrict__ stop) {
  if (stop && *stop) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Cref[i] = C[i];
  if (i == 0) *dmax = 0.0f;
}

// home[w] = the label (from the previous assign) of the first point of
// workgroup w's slice of the order, clamped into [0, K)
__global__ void k_wg_home(long nwg, long M, const int *__restrict__ Mdev, int K,
                          const int *__restrict__ order, const int *__restrict__ labels,
                          int *__restrict__ home) {
  const long w = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwg) return;
  if (Mdev) M = min(M, (long)*Mdev);  // the open points only (drift bounds)
  if (w * MF_PTS >= M) {  // a workgroup with no points (it exits at once)
    home[w] = 0;
    return;
  }
  const long q = w * MF_PTS;
  const int l = labels[order[q]];
  home[w] = l < 0 ? 0 : (l >= K ? K - 1 : l);
}

// top-3 of packed values: b1 >= b2 >= b3
// (max written as med3 with a run-time +FLT_MAX the compiler cannot fold:
// one instruction, no NaN canonicalisation of the bit-packed input)
__device__ inline void top3_insert(float p, float &b1, float &b2, float &b3, float big) {
  const float n3 = __builtin_amdgcn_fmed3f(b2, b3, p);
  const float n2 = __builtin_amdgcn_fmed3f(b1, b2, p);
  b1 = __builtin_amdgcn_fmed3f(b1, p, big);
  b2 = n2;
  b3 = n3;
}

struct Top3 {
  float b1 = -FLT_MAX, b2 = -FLT_MAX, b3 = -FLT_MAX;
  int t1 = 0, t2 = 0;
  // label-ordered assign: U = the largest score of the tiles whose insertion
  // was skipped, th = the skip threshold, mg = 2E in MFMA score units; hs =
  // the point's approximate score of the workgroup's home centre (MFMA
  // units), xn >= |x|, E the error bound (FP32 units)
  float U = -INFINITY, th = -INFINITY, mg = 0.0f;
  float hs = -INFINITY, xn = 0.0f, E = 0.0f;
};

// Forward-error bound E (FP32 score units) of every centre's approximate
// score of one point: xn >= |x|, ex >= |x - xh/s|, the centre bounds cmax,
// ncmax, dcmax, eta the FP16 subnormal margin (0 for BF16).  The terms are
// those of the certificate described at the top of this file.
__host__ __device__ inline float mf_err_bound(float xn, float ex, float Df, float cmax,
                                              float ncmax, float dcmax, float eta) {
  const float g32 = Df * 5.9604645e-8f * 1.0001f;         // gamma_D, u = 2^-24
  const float gmf = (Df + 1.0f) * 2.3841858e-7f * 1.001f;  // gamma'_{D+1}, u = 2^-22
  const float rnd = ex * cmax + (xn + ex) * dcmax;
  // FP16 engines: margin for subnormal operands the matrix core may flush
  const float etat = eta * sqrtf(Df) * (xn + cmax) + eta * eta * Df;
  const float mfe = gmf * (0.5f * ncmax + (xn + ex) * (cmax + dcmax) + etat);
  const float fpe = g32 * (0.5f * ncmax + xn * cmax);
  return (rnd + etat + mfe + fpe) * 1.001f + 1e-30f;
}

// row of output register r (0..15) of a 32x32 MFMA tile for lane half h
__device__ inline int mf_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// Label-ordered assign helpers (ORD): the centre tile visited at step kt
// (Ts = the workgroup's tile order in LDS) and the point at sorted position q.
// (The DMA addresses in k_assign_mfma are written out inline: with a call in
// a global_load_lds address argument, the host compilation of this file
// drops the kernel instantiations without a diagnostic.)
template <bool ORD>
__host__ __device__ __forceinline__ int mf_tile(const int *Ts, int kt) {
  return ORD ? Ts[kt] : kt;
}
template <bool ORD>
__host__ __device__ __forceinline__ long mf_point(const int *order, long q) {
  return ORD ? (long)order[q] : q;
}

// ORD: the workgroups take the points in the order `order` (the previous
// update's stable sort by label, so a workgroup's points mostly share their
// previous label) and visit the centre tiles in the order ttab[home[wg]]
// (tiles nearest to the centre of that label first).  A point's top three
// then settle within the first few tiles, and each later tile's insertion is
// skipped when no lane of the wave holds a score that could enter its top
// three (the test is exact, so the state is the one a full insertion leaves).
// The certificates do not depend on the order in which centres are visited,
// so the labels are those of the unordered kernel.
template <int KS, bool FP16, bool ORD, bool LIST>
__global__ __launch_bounds__(64 * MF_WAVES) void k_assign_mfma(
    long M, const uint16_t *__restrict__ Xh, const float *__restrict__ X, int D, int K, int Kp,
    const uint16_t *__restrict__ Ch, const float *__restrict__ bias, const float *__restrict__ C,
    const float *__restrict__ nc, const float *__restrict__ bnd, const float2 *__restrict__ xnb,
    float inv_s2, float eta, float big, int *__restrict__ labels,
    int *__restrict__ flags, int *__restrict__ nflag, int *__restrict__ ncand,
    const int *__restrict__ stop, const int *__restrict__ order, const int *__restrict__ ttab,
    const float *__restrict__ tdist, const int *__restrict__ home, const int *__restrict__ Mdev,
    float4 *__restrict__ xbo) {
  // stop (nullable): set by a device-side convergence test; the launch then
  // assigns nothing and lists nothing
  if (stop && *stop) return;
  // Mdev (nullable): the number of positions of `order` to assign (the points
  // the drift bounds left open; the others keep their labels), counted into
  // slot 3 of the first counter line; xbo (nullable): fresh drift bounds of
  // the assigned points
  if (Mdev) {
    const long Mh = M;
    M = min(M, (long)*Mdev);
    if (blockIdx.x == 0 && threadIdx.x == 0) ncand[3] = (int)(Mh - M);
  }
  // the points this launch scores, into slot 5 (one launch per tile of
  // streamed points adds into the same iteration's counters)
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(ncand + 5, (int)M);
  if (Mdev && (long)blockIdx.x * MF_PTS >= M) return;  // uniform over the workgroup
  constexpr int Dp = KS * 16;
  constexpr int CPR = Dp / 8;               // 16-byte chunks per centre row
  constexpr int SROWS = MF_STAGE * MF_TILE; // centre rows per LDS stage
  constexpr int SCH = SROWS * CPR;          // chunks per stage
  constexpr int ROTM = CPR >= 16 ? 15 : 7;   // row rotation mask of the LDS swizzle
  // LDS: two stages of centre rows (chunk q of row r stored at position
  // (q + r) % CPR, so the 32 rows one ds_read_b128 touches fall in distinct
  // banks) and two stages of biases; filled by global->LDS DMA.
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  u32x4_t *As = reinterpret_cast<u32x4_t *>(smem);                         // [2][SCH]
  float *Bs = reinterpret_cast<float *>(smem + 2 * SCH * sizeof(u32x4_t));  // [2][SROWS]
  int *Ts = reinterpret_cast<int *>(Bs + 2 * SROWS);                        // [nKT] (ORD)
  float *Dt = reinterpret_cast<float *>(Ts + Kp / MF_TILE);                 // [nKT] (ORD)
  __shared__ int cutv[2][MF_WAVES];                                         // (ORD)
  typedef typename std::conditional<FP16, f16x8_t, bf16x8_t>::type vec8;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, col = lane & 31;
  const long p0 = (long)blockIdx.x * MF_PTS + w * 64;  // sorted positions of the wave
  const int nKT = Kp / MF_TILE;      // multiple of MF_KPAD / MF_TILE = 4
  const int nS = nKT / MF_STAGE;
  // ORD: the home centre, its tile and the register (hr) and lane half (hh)
  // that hold its row of that tile's scores
  const int hc = ORD ? home[blockIdx.x] : 0;
  const int htile = hc / MF_TILE, hrow = hc % MF_TILE;
  const int hh = (hrow >> 2) & 1, hr = ((hrow >> 3) << 2) | (hrow & 3);
  // point fragments: lane holds the point at sorted position p0 + g*32 + col,
  // k = ks*16 + 8h + 0..7 (issued first: the loads below do not wait on them)
  u32x4_t bfrag[2][KS];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const long sp = p0 + g * 32 + col;
    const long p = sp < M ? mf_point<ORD || LIST>(order, sp) : 0;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (sp < M)
        bfrag[g][ks] = *reinterpret_cast<const u32x4_t *>(Xh + p * Dp + ks * 16 + 8 * h);
      else
        bfrag[g][ks] = u32x4_t{0u, 0u, 0u, 0u};
    }
  }
  int tsv = 0;  // ORD: this thread's entries of the tile order and bounds
  float dsv = 0.0f;
  if constexpr (ORD) {
    if (tid < nKT) {
      tsv = ttab[(long)hc * nKT + tid];
      dsv = tdist[(long)hc * nKT + tid];
    }
  }
  if constexpr (ORD) {  // the workgroup's tile order and distance bounds
    for (int t = tid; t < nKT; t += 64 * MF_WAVES) {
      Ts[t] = t == tid ? tsv : ttab[(long)hc * nKT + t];
      Dt[t] = t == tid ? dsv : tdist[(long)hc * nKT + t];
    }
    __syncthreads();
  }

  // per point group: the three largest packed scores and the tiles of the
  // first two (named structs, not arrays: selects between array elements
  // would force the state into scratch memory)
  Top3 S0, S1;
  if constexpr (ORD) {  // 2E of each lane's point, in MFMA score units
    const float cmx = bnd[0], ncx = bnd[1], dcx = bnd[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const long sp = p0 + g * 32 + col;
      const float2 nb = sp < M ? xnb[mf_point<ORD>(order, sp)] : make_float2(0.0f, 0.0f);
      const float E = mf_err_bound(nb.x, nb.y, (float)D, cmx, ncx, dcx, eta);
      Top3 &T = g == 0 ? S0 : S1;
      T.mg = 2.0f * E / inv_s2;
      T.E = E;
      T.xn = nb.x;
    }
  }

  // asynchronous copy of centre stage st into LDS buffer b: each DMA
  // instruction writes 64 consecutive chunks (one per lane, in lane order).
  // Chunk q of row r lives at position (q + rot(r)) % CPR with rot(r) = r & ROTM
  // (15 for rows of 16 or more chunks, else 7)
  // (for CPR >= 8), so the 8 rows one ds_read_b128 phase touches hit
  // distinct banks.
  auto issue_stage = [&](int st, int b) {
    u32x4_t *dst = As + b * SCH;
    constexpr int NI = SCH / 64;                            // DMA instructions per stage
    constexpr int PERW = (NI + MF_WAVES - 1) / MF_WAVES;    // per wave (the last may idle)
    if constexpr (64 % CPR == 0) {
      constexpr int RPI = 64 / CPR;  // rows per DMA instruction
#pragma unroll
      for (int i = 0; i < PERW; ++i) {
        const int n = i * MF_WAVES + w;
        if (n < NI) {
          const int r0 = n * RPI + lane / CPR;
          const int rot = CPR >= 8 ? (r0 & ROTM) : 0;
          const int q = (lane % CPR - rot + CPR) % CPR;
          const int tl = st * MF_STAGE + r0 / MF_TILE;
          const long row = (long)(ORD ? Ts[tl] : tl) * MF_TILE + (r0 % MF_TILE);
          __builtin_amdgcn_global_load_lds(Ch + row * Dp + q * 8, dst + n * 64, 16, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PERW; ++i) {
        const int n = i * MF_WAVES + w;
        if (n < NI) {
          const int c0 = n * 64;
          const int c = c0 + lane;
          const int r = c / CPR, pos = c - r * CPR;
          const int rot = CPR >= 8 ? (r & ROTM) : 0;
          const int q = (pos - rot + CPR) % CPR;
          const int tl = st * MF_STAGE + r / MF_TILE;
          const long row = (long)(ORD ? Ts[tl] : tl) * MF_TILE + (r % MF_TILE);
          __builtin_amdgcn_global_load_lds(Ch + row * Dp + q * 8, dst + c0, 16, 0, 0);
        }
      }
    }
    if (tid < SROWS / 4)  // biases: 4 floats per lane
    {
      const int tl = st * MF_STAGE + (tid * 4) / MF_TILE;
      const long row = (long)(ORD ? Ts[tl] : tl) * MF_TILE + ((tid * 4) % MF_TILE);
      __builtin_amdgcn_global_load_lds(bias + row, Bs + b * SROWS, 16, 0, 0);
    }
  };

  // scores of one centre tile for both point groups; bias is the
  // accumulator input (rows 8*(r/4) + 4h + r%4)
  auto mfma_tile = [&](int kt, f32x16_t &a0, f32x16_t &a1) {
    const int b = (kt / MF_STAGE) & 1, j = kt % MF_STAGE;
    const u32x4_t *A = As + b * SCH;
    const float *Bb = Bs + b * SROWS + j * MF_TILE;
    f32x16_t cinit;
#pragma unroll
    for (int r = 0; r < 16; ++r) cinit[r] = Bb[mf_row(r, h)];
    a0 = cinit;
    a1 = cinit;
    const int row = j * MF_TILE + col;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int q = ks * 2 + h;
      const u32x4_t a = A[row * CPR + (q + (CPR >= 8 ? (row & ROTM) : 0)) % CPR];
      if constexpr (FP16) {
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vec8, a),
                                                    __builtin_bit_cast(vec8, bfrag[0][ks]), a0, 0,
                                                    0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vec8, a),
                                                    __builtin_bit_cast(vec8, bfrag[1][ks]), a1, 0,
                                                    0, 0);
      } else {
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vec8, a),
                                                     __builtin_bit_cast(vec8, bfrag[0][ks]), a0, 0,
                                                     0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vec8, a),
                                                     __builtin_bit_cast(vec8, bfrag[1][ks]), a1, 0,
                                                     0, 0);
      }
    }
  };
  // pack the register id into the low 4 bits, keep the top 3 and the tiles
  // of the first two.  ORD: a tile whose every score lies more than 2E (plus
  // a slack) below the point's best so far (both lane halves) cannot hold a
  // competitor of it; when that holds for every lane of the wave the
  // insertion is skipped and only the tile's largest score is kept, in U.
  // The top three are then exact over the tiles inserted and U bounds every
  // other score, which is all the certificates use (see finish).
  auto epi_group = [&](int kt, const f32x16_t &acc, Top3 &T) {
    if constexpr (ORD) {
      if (mf_tile<ORD>(Ts, kt) == htile) {  // the home centre's score of each point
        float v = acc[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) v = r == hr ? acc[r] : v;
        const float o = __shfl_xor(v, 32);
        T.hs = h == hh ? v : o;
      }
      float m = acc[0];
#pragma unroll
      for (int r = 1; r < 16; ++r) m = __builtin_fmaxf(m, acc[r]);
      if (__ballot(!(m < T.th)) == 0) {
        T.U = __builtin_fmaxf(T.U, m);
        return;
      }
    }
    kt = mf_tile<ORD>(Ts, kt);
    const float o1 = T.b1, o2 = T.b2;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float p = __uint_as_float((__float_as_uint(acc[r]) & 0xFFFFFFF0u) | (unsigned)r);
      top3_insert(p, T.b1, T.b2, T.b3, big);
    }
    const unsigned u1 = __float_as_uint(T.b1), u2 = __float_as_uint(T.b2);
    const int ot1 = T.t1, ot2 = T.t2;
    const int nt2 = (u2 == __float_as_uint(o1)) ? ot1 : (u2 == __float_as_uint(o2)) ? ot2 : kt;
    T.t1 = (u1 != __float_as_uint(o1)) ? kt : ot1;
    T.t2 = nt2;
    if constexpr (ORD) {
      const float bc = __builtin_fmaxf(T.b1, __shfl_xor(T.b1, 32));
      T.th = bc - T.mg - __builtin_fabsf(bc) * 1e-4f;
    }
  };
  auto epi_tile = [&](int kt, const f32x16_t &a0, const f32x16_t &a1) {
    epi_group(kt, a0, S0);
    epi_group(kt, a1, S1);
  };

  // Software pipeline: while the VALU runs the epilogue of tile t, the
  // matrix core computes tile t+1 (same basic block); stage s+1 streams into
  // the other LDS buffer while stage s is consumed.  nKT is a multiple of
  // 4 (and so of MF_STAGE), so the loop covers tiles [0, nKT - 2) and the
  // last pair is peeled: no tile past the end is computed and no barrier is
  // waited at after the last stage.
  // ORD cut: every centre k of a tile at rank >= kn is at least Dt[kn] from
  // the home centre c, so |x - c_k| >= Dt[kn] - |x - c| (triangle
  // inequality) and its exact score x.c_k - 0.5|c_k|^2 = 0.5 (|x|^2 -
  // |x - c_k|^2) is at most 0.5 (xn^2 - g^2) with g = Dt[kn] - hb, hb >=
  // |x - c| from the home score (exact >= hs - E): hb^2 = xn^2 - 2 (hs - E).
  // That plus E bounds the approximate scores of all those tiles.  When the
  // bound is below the skip threshold for every point of the workgroup, the
  // remaining tiles are not computed and the bound joins U, as for a
  // skipped insertion.  (Roundings of the bound are covered by relative
  // margins of 1e-6 and more; Dt is itself a lower bound.)
  auto cut_bound = [&](const Top3 &T, float dn, float &ubm) -> bool {
    const float a_h = T.hs * inv_s2;
    const float x2 = T.xn * T.xn * (1.0f + 2.4e-7f);
    const float h2 = (x2 - 2.0f * (a_h - T.E)) +
                     3.6e-7f * (x2 + 2.0f * __builtin_fabsf(a_h) + 2.0f * T.E);
    const float hb = __builtin_sqrtf(__builtin_fmaxf(h2, 0.0f)) * (1.0f + 1e-6f);
    const float g = (dn - hb) * (1.0f - 1e-6f);
    float ub = 0.5f * (x2 - g * g) + T.E;
    ub += 1e-6f * (x2 + g * g + T.E);
    ubm = ub / inv_s2;
    ubm += 1e-6f * __builtin_fabsf(ubm);
    return g > 0.0f && ubm < T.th;
  };
  issue_stage(0, 0);
  __builtin_amdgcn_s_waitcnt(0);  // point fragments and stage 0 have arrived
  __syncthreads();
  f32x16_t A0, A1, B0, B1;
  mfma_tile(0, A0, A1);
  int kt = 0;
  bool cut = false;
  for (; kt + 2 < nKT; kt += 2) {
    if (kt % MF_STAGE == 0 && kt / MF_STAGE + 1 < nS) issue_stage(kt / MF_STAGE + 1, (kt / MF_STAGE + 1) & 1);
    mfma_tile(kt + 1, B0, B1);
    epi_tile(kt, A0, A1);
    if ((kt + 2) % MF_STAGE == 0) {  // next tile opens a new stage
      __builtin_amdgcn_s_waitcnt(0);
      if constexpr (ORD) {
        // workgroup vote: each wave's verdict in LDS, double-buffered by
        // stage parity (a wave can be at most one stage boundary ahead)
        const float dn = Dt[kt + 2];
        float u0 = -INFINITY, u1 = -INFINITY;
        const bool ok0 = !(p0 + col < M) || cut_bound(S0, dn, u0);
        const bool ok1 = !(p0 + 32 + col < M) || cut_bound(S1, dn, u1);
        const int par = ((kt + 2) / MF_STAGE) & 1;
        const bool wok = __ballot(!(ok0 && ok1)) == 0;
        if (lane == 0) cutv[par][w] = wok ? 1 : 0;
        __syncthreads();
        int all = 1;
#pragma unroll
        for (int u = 0; u < MF_WAVES; ++u) all &= cutv[par][u];
        if (all) {
          epi_tile(kt + 1, B0, B1);
          S0.U = __builtin_fmaxf(S0.U, u0);
          S1.U = __builtin_fmaxf(S1.U, u1);
          cut = true;
          break;
        }
      } else {
        __syncthreads();
      }
    }
    mfma_tile(kt + 2, A0, A1);
    epi_tile(kt + 1, B0, B1);
  }
  if (!cut) {
    mfma_tile(kt + 1, B0, B1);  // kt = nKT - 2: a pair within the last stage
    epi_tile(kt, A0, A1);
    epi_tile(kt + 1, B0, B1);
  }
  // tiles computed and tiles in all, beside the rescored count of the same
  // counter line
  if (tid == 0) {
    int *cl = ncand + (blockIdx.x & (NC_SPREAD - 1)) * NC_STRIDE;
    atomicAdd(cl + 1, cut ? kt + 2 : nKT);
    atomicAdd(cl + 2, nKT);
  }

  // merge the two lane halves (same points, different centre rows)
  const float cmax = bnd[0], ncmax = bnd[1], dcmax = bnd[2];
  auto finish = [&](int g, const Top3 &T) {
    // candidate indices of this half
    const unsigned u1 = __float_as_uint(T.b1), u2 = __float_as_uint(T.b2);
    const int i1 = T.t1 * MF_TILE + mf_row((int)(u1 & 15u), h);
    const int i2 = T.t2 * MF_TILE + mf_row((int)(u2 & 15u), h);
    bool tie = (u1 == u2) || (u2 == __float_as_uint(T.b3));
    const float v1 = T.b1, v2 = T.b2, v3 = T.b3;
    // the other half's triple
    const float w1 = __shfl_xor(v1, 32), w2 = __shfl_xor(v2, 32), w3 = __shfl_xor(v3, 32);
    const int j1 = __shfl_xor(i1, 32), j2 = __shfl_xor(i2, 32);
    const bool otie = __shfl_xor((int)tie, 32) != 0;
    // merge two sorted triples (exact bitwise duplicates are flagged); the
    // two lists are disjoint sets of centres, each sorted descending
    float m1, m2, m3;
    int k1, k2;
    {
      const bool A1 = v1 >= w1;                  // first from this half?
      m1 = A1 ? v1 : w1;
      k1 = A1 ? i1 : j1;
      const float x2 = A1 ? v2 : v1, y2 = A1 ? w1 : w2;  // heads after taking one
      const int xi2 = A1 ? i2 : i1, yi2 = A1 ? j1 : j2;
      const bool A2 = x2 >= y2;
      m2 = A2 ? x2 : y2;
      k2 = A2 ? xi2 : yi2;
      // third: the larger of the two remaining heads
      const float xa = A1 ? (A2 ? v3 : v2) : (A2 ? v2 : v1);
      const float ya = A1 ? (A2 ? w1 : w2) : (A2 ? w2 : w3);
      m3 = fmaxf(xa, ya);
      tie = tie || otie || (__float_as_uint(m1) == __float_as_uint(m2)) ||
            (__float_as_uint(m2) == __float_as_uint(m3));
    }
    const long sp = p0 + g * 32 + col;
    const bool inr = sp < M;
    const long p = inr ? mf_point<ORD || LIST>(order, sp) : 0;
    if constexpr (ORD) {  // the skipped tiles' scores are at most Ub (both halves)
      const float Ub = fmaxf(T.U, __shfl_xor(T.U, 32));
      m2 = fmaxf(m2, Ub);
      m3 = fmaxf(m3, Ub);
    }
    // forward-error bound E for every centre's approximate score (FP32 units)
    const float2 nb = inr ? xnb[p] : make_float2(0.0f, 0.0f);
    const float E = mf_err_bound(nb.x, nb.y, (float)D, cmax, ncmax, dcmax, eta);
    const float a1 = m1 * inv_s2, a2 = m2 * inv_s2, a3 = m3 * inv_s2;  // FP32 score units
    // packed values carry 4 replaced low bits (< 16 ulp); FP32 subtraction
    // and the scaling add a few more ulps
    const float slack = (fabsf(a1) + fabsf(a2) + fabsf(a3)) * 1.6e-5f;
    const bool live = h == 0 && inr;
    // certificate 1: the best approximate score beats every other by > 2E
    const bool quick = !tie && (a1 - a2 > 2.0f * E + slack) && k1 >= 0 && k1 < K;
    // certificate 2: the exact FP32 chains (the plain path's arithmetic) of
    // the two candidates, computed here by the two lane halves (k1 by half
    // 0, k2 by half 1); the better one (lowest index on equal scores) must
    // beat every other centre's upper bound a3 + E; with k <= 2 there is no
    // other centre
    const bool has_third = m3 > -FLT_MAX * 0.5f;
    const float thr = has_third ? a3 + E + slack : -INFINITY;
    const bool need = inr && !quick && !tie && k1 >= 0 && k1 < K;
    float sx = -INFINITY;
    if (need) {
      const int cj = h == 0 ? k1 : k2;
      if (cj >= 0 && cj < K) {
        const float *xr = X + p * D;
        const float *cr = C + (long)cj * D;
        float sc = -0.5f * nc[cj];
        if ((D & 3) == 0 && ((uintptr_t)X & 15) == 0 && ((uintptr_t)C & 15) == 0) {
          // 16-byte pieces, 16 of each row in flight at a time (the FP32 row
          // of the point comes from HBM), added in ascending d
          const float4 *x4 = reinterpret_cast<const float4 *>(xr);
          const float4 *c4 = reinterpret_cast<const float4 *>(cr);
          const int n4 = D / 4;
          for (int b0 = 0; b0 < n4; b0 += 16) {
            float4 xv[16], cv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
              if (b0 + u < n4) {
                xv[u] = x4[b0 + u];
                cv[u] = c4[b0 + u];
              }
#pragma unroll
            for (int u = 0; u < 16; ++u)
              if (b0 + u < n4) {
                sc = __builtin_fmaf(xv[u].x, cv[u].x, sc);
                sc = __builtin_fmaf(xv[u].y, cv[u].y, sc);
                sc = __builtin_fmaf(xv[u].z, cv[u].z, sc);
                sc = __builtin_fmaf(xv[u].w, cv[u].w, sc);
              }
          }
        } else {
          for (int q = 0; q < D; ++q) sc = __builtin_fmaf(xr[q], cr[q], sc);
        }
        sx = sc;
      }
    }
    const float sy = __shfl_xor(sx, 32);  // the other half's chain
    bool fall2 = false;
    float sb = sx, so = sy;  // the chains of the label and of the other candidate
    int lab = k1, oth = k2;
    if (live && need) {
      if (k2 >= 0 && k2 < K && (sy > sx || (sy == sx && k2 < k1))) {
        lab = k2;
        oth = k1;
        sb = sy;
        so = sx;
      }
      if (sb > thr && sb > -INFINITY) labels[p] = lab;
      else fall2 = true;
    }
    const bool two = live && need;
    const bool fall = (live && !quick && !need) || fall2;
    if (live && quick) labels[p] = k1;
    if (xbo && live) {
      // fresh drift bounds: lo <= the exact score of the label (the
      // approximate one minus E and the packing slack, or its FP32 chain
      // minus E >= the chain's own bound), hi >= the exact score of every
      // other centre, hi3 >= that of every centre but the label and the
      // runner-up rr (the other candidate: every centre but the two has an
      // approximate score at most a3); distances from |x - c|^2 = |x|^2 - 2
      // score with the upper and lower bounds of |x|^2 that xn gives.  A
      // point sent on to the later passes gets bounds that never hold.
      float lo = -INFINITY, hi = INFINITY, hi3 = INFINITY;
      int rr = -1, lb = k1;
      if (quick) {
        lo = a1 - E - slack;
        hi = a2 + E + slack;
        rr = k2;
      } else if (need && !fall2) {
        lo = sb - E;
        hi = fmaxf(a3 + E + slack, so + E);
        rr = oth;
        lb = lab;
      }
      if (lo > -INFINITY) hi3 = has_third ? a3 + E + slack : -INFINITY;
      if (rr < 0 || rr >= K || rr == lb) rr = -1;
      float u = INFINITY, l = 0.0f, l3 = 0.0f;
      if (lo > -INFINITY) {
        u = dist_upper(nb.x, lo);
        l = dist_lower(nb.x, hi, (float)D);
        l3 = hi3 == -INFINITY ? INFINITY : dist_lower(nb.x, hi3, (float)D);
      }
      xbo[p] = make_float4(u, l, l3, __int_as_float(rr));
    }
    // wave-aggregated counts and appends (one atomic per wave)
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    {
      const uint64_t mk = __ballot(two);
      if (mk && lane == __ffsll((long long)mk) - 1)
        atomicAdd(ncand + (blockIdx.x & (NC_SPREAD - 1)) * NC_STRIDE, __popcll(mk));
    }
    {
      const uint64_t mk = __ballot(fall);
      if (mk) {
        const int lead = __ffsll((long long)mk) - 1;
        int base = 0;
        if (lane == lead) base = atomicAdd(nflag, __popcll(mk));
        base = __shfl(base, lead);
        if (fall) flags[base + __popcll(mk & below)] = (int)p;
      }
    }
  };
  finish(0, S0);
  finish(1, S1);
}

// Deep re-resolve of the points certificate 2 left open (typically many
// centres within the BF16 error bound of the best).  A second MFMA pass over
// just those points in split BF16: x = xh + xl and c = ch + cl (the low parts
// are the BF16 roundings of the residuals), scores xh.ch + xh.cl + xl.ch by
// three MFMAs, whose operand error is at most u^2 (3 + 2u) sum|x_d c_d|
// instead of (2u + u^2) sum|x_d c_d|.  Each lane half keeps the four largest
// approximate scores of its centre rows, packed with the register id in the
// low 4 bits as in k_assign_mfma (four v_med3 per score), and the tiles of
// the first three; the halves' lists are merged (each reads the other's
// original list), and two equal packed values or a repeated candidate among
// the merged four send the point on, since their tiles cannot be told apart.
// The pass computes the exact FP32 chains of the first three (half 0 of the
// wave two, half 1 one, each reading its point's row once) and labels the
// point when the best of them (lowest index on equal scores) exceeds the
// fourth + E + the packing slack; the rest go to the FP32 full search.  One
// wave handles 32 listed points; centre tiles and their biases are staged
// in LDS.  Both engines: the parts are unscaled BF16 (for FP16 engines the
// BF16 centre copies Chb/Clb and biases biasb), computed from the FP32 rows.
template <int KS>
__global__ __launch_bounds__(256) void k_deep_mfma(
    const int *__restrict__ nlist, const int *__restrict__ list,
    const float *__restrict__ X, int D, int K, int Kp, const uint16_t *__restrict__ Ch,
    const uint16_t *__restrict__ Cl, const float *__restrict__ bias, const float *__restrict__ C,
    const float *__restrict__ nc, const float *__restrict__ bnd, const float2 *__restrict__ xnb,
    float big, int *__restrict__ labels, int *__restrict__ list2, int *__restrict__ nlist2,
    float4 *__restrict__ xbo) {
  // xbo (nullable): fresh drift bounds of the points this pass labels
  const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const long total = *nlist;
  const long wstride = (long)gridDim.x * 128;
  const float cmax = bnd[0], ncmax = bnd[1];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  constexpr int Dp = KS * 16;
  constexpr int CH = MF_TILE * Dp / 8;  // 16-byte chunks of one tile (high or low part)
  constexpr int PER = (2 * CH + 255) / 256;
  // two LDS buffers of one centre tile: high chunks then low chunks; chunk q
  // of row r at r*CPR + (q + r) % CPR (rotation: the 32 rows one fragment
  // read touches fall in different banks)
  constexpr int CPR = Dp / 8;
  __shared__ u32x4_t ts[2][2 * CH];
  __shared__ float tb[2][MF_TILE];  // the tile's biases
  const int nT = Kp / MF_TILE;
  auto fetch = [&](int t, u32x4_t *reg, float &rb) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int c = threadIdx.x + 256 * i;
      if (c < 2 * CH) {
        const int part = c >= CH, cc = part ? c - CH : c;
        const uint16_t *src = part ? Cl : Ch;
        reg[i] = *reinterpret_cast<const u32x4_t *>(src + (long)t * MF_TILE * Dp + (long)cc * 8);
      }
    }
    if (threadIdx.x < MF_TILE) rb = bias[t * MF_TILE + threadIdx.x];
  };
  // the loop over point groups is uniform across the workgroup (barriers)
  for (long bb = (long)blockIdx.x * 128; bb < total; bb += wstride) {
    const long b0 = bb + (threadIdx.x >> 6) * 32;
    const long li = b0 + col;
    const bool ok = li < total;
    const long p = ok ? (long)list[li] : 0;
    // point fragments (8 consecutive dimensions per k-step): high and low
    // BF16 parts of the FP32 row (v - high is exact in FP32)
    u32x4_t xh4[KS], xl4[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      xh4[ks] = u32x4_t{0u, 0u, 0u, 0u};
      xl4[ks] = u32x4_t{0u, 0u, 0u, 0u};
      if (ok) {
        // this half's 8 dimensions of the k-step: two 16-byte loads when they
        // all lie inside a row whose length is a multiple of 8 (16-byte
        // aligned), else one guarded load each
        const int db = ks * 16 + 8 * h;
        float xs[8];
        if ((D & 7) == 0 && db + 8 <= D && ((uintptr_t)X & 15) == 0) {
          const float4 lo4 = *reinterpret_cast<const float4 *>(X + p * D + db);
          const float4 hi4 = *reinterpret_cast<const float4 *>(X + p * D + db + 4);
          xs[0] = lo4.x; xs[1] = lo4.y; xs[2] = lo4.z; xs[3] = lo4.w;
          xs[4] = hi4.x; xs[5] = hi4.y; xs[6] = hi4.z; xs[7] = hi4.w;
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) xs[u] = db + u < D ? X[p * D + db + u] : 0.0f;
        }
        uint32_t wh[4], wl[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float v0 = xs[2 * u];
          const float v1 = xs[2 * u + 1];
          const uint16_t h0 = f32_to_bf16_rne(v0), h1 = f32_to_bf16_rne(v1);
          const float r0 = v0 - __uint_as_float((uint32_t)h0 << 16);
          const float r1 = v1 - __uint_as_float((uint32_t)h1 << 16);
          wh[u] = (uint32_t)h0 | ((uint32_t)h1 << 16);
          wl[u] = (uint32_t)f32_to_bf16_rne(r0) | ((uint32_t)f32_to_bf16_rne(r1) << 16);
        }
        xh4[ks] = u32x4_t{wh[0], wh[1], wh[2], wh[3]};
        xl4[ks] = u32x4_t{wl[0], wl[1], wl[2], wl[3]};
      }
    }
    // per lane half: the four largest packed scores (register id in the low
    // 4 bits, as in k_assign_mfma) and the tiles of the first three
    float b1 = -FLT_MAX, b2 = -FLT_MAX, b3 = -FLT_MAX, b4 = -FLT_MAX;
    int t1 = 0, t2 = 0, t3 = 0;
    u32x4_t pre[PER];
    float preb = 0.0f;
    fetch(0, pre, preb);
    for (int t = 0; t < nT; ++t) {
      const int buf = t & 1;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int c = threadIdx.x + 256 * i;
        if (c < 2 * CH) {
          const int part = c >= CH, cc = part ? c - CH : c;
          const int r = cc / CPR, q = cc - r * CPR;
          ts[buf][part * CH + r * CPR + (q + r) % CPR] = pre[i];
        }
      }
      if (threadIdx.x < MF_TILE) tb[buf][threadIdx.x] = preb;
      __syncthreads();
      if (t + 1 < nT) fetch(t + 1, pre, preb);
      f32x16_t acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = tb[buf][mf_row(r, h)];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int q = col * CPR + (ks * 2 + h + col) % CPR;
        const u32x4_t ah = ts[buf][q];
        const u32x4_t al = ts[buf][CH + q];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, ah),
                                                       __builtin_bit_cast(bf16x8_t, xh4[ks]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, ah),
                                                       __builtin_bit_cast(bf16x8_t, xl4[ks]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, al),
                                                       __builtin_bit_cast(bf16x8_t, xh4[ks]), acc, 0, 0, 0);
      }
      const unsigned o1 = __float_as_uint(b1), o2 = __float_as_uint(b2), o3 = __float_as_uint(b3);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float pv = __uint_as_float((__float_as_uint(acc[r]) & 0xFFFFFFF0u) | (unsigned)r);
        const float n4 = __builtin_amdgcn_fmed3f(b3, b4, pv);
        const float n3 = __builtin_amdgcn_fmed3f(b2, b3, pv);
        const float n2 = __builtin_amdgcn_fmed3f(b1, b2, pv);
        b1 = __builtin_amdgcn_fmed3f(b1, pv, big);
        b2 = n2;
        b3 = n3;
        b4 = n4;
      }
      // an entry is an earlier entry (entries only move down) or from tile t;
      // bitwise-equal entries make this ambiguous and are flagged below
      const unsigned u1 = __float_as_uint(b1), u2 = __float_as_uint(b2), u3 = __float_as_uint(b3);
      const int n1 = u1 == o1 ? t1 : t;
      const int n2 = u2 == o1 ? t1 : u2 == o2 ? t2 : t;
      const int n3 = u3 == o1 ? t1 : u3 == o2 ? t2 : u3 == o3 ? t3 : t;
      t1 = n1;
      t2 = n2;
      t3 = n3;
    }
    // unpacked lists: the first three with their centres, the fourth (value
    // only) as the bound
    float bv[4] = {b1, b2, b3, b4};
    int bi[4] = {t1 * MF_TILE + mf_row((int)(__float_as_uint(b1) & 15u), h),
                 t2 * MF_TILE + mf_row((int)(__float_as_uint(b2) & 15u), h),
                 t3 * MF_TILE + mf_row((int)(__float_as_uint(b3) & 15u), h), -1};
    // merge the other half's four (disjoint centres), read before any update
    float ov[4];
    int oi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ov[q] = __shfl_xor(bv[q], 32);
      oi[q] = __shfl_xor(bi[q], 32);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float v = ov[q];
      int vi = oi[q];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (v > bv[u]) {
          const float tv = bv[u];
          const int ti = bi[u];
          bv[u] = v; bi[u] = vi; v = tv; vi = ti;
        }
      }
    }
    // two equal packed values among the four (their tiles cannot be told
    // apart) or a repeated candidate send the point on
    const float lo = -FLT_MAX * 0.5f;
    const bool tie = (bv[1] > lo && __float_as_uint(bv[0]) == __float_as_uint(bv[1])) ||
                     (bv[2] > lo && __float_as_uint(bv[1]) == __float_as_uint(bv[2])) ||
                     (bv[3] > lo && __float_as_uint(bv[2]) == __float_as_uint(bv[3])) ||
                     bi[0] == bi[1] || bi[1] == bi[2] || bi[0] == bi[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bv[q] = __shfl(bv[q], col);
      bi[q] = __shfl(bi[q], col);
    }
    // exact chains, one pass over the row: half 0 -> entries 0, 1; half 1 -> 2
    const int c0 = h ? bi[2] : bi[0];
    const int c1 = h ? -1 : bi[1];
    const bool l0 = ok && c0 >= 0 && c0 < K, l1 = ok && c1 >= 0 && c1 < K;
    float s0 = l0 ? -0.5f * nc[c0] : -INFINITY, s1 = l1 ? -0.5f * nc[c1] : -INFINITY;
    if (l0 || l1) {
      const float *xr = X + p * D;
      const float *r0 = C + (long)(l0 ? c0 : 0) * D, *r1 = C + (long)(l1 ? c1 : 0) * D;
      if ((D & 3) == 0 && (((uintptr_t)X | (uintptr_t)C) & 15) == 0) {
        const float4 *x4 = reinterpret_cast<const float4 *>(xr);
        const float4 *a4 = reinterpret_cast<const float4 *>(r0);
        const float4 *b4 = reinterpret_cast<const float4 *>(r1);
#pragma unroll 4
        for (int q = 0; q < D / 4; ++q) {
          const float4 xv = x4[q], av = a4[q], bw = b4[q];
          s0 = __builtin_fmaf(xv.x, av.x, s0);
          s1 = __builtin_fmaf(xv.x, bw.x, s1);
          s0 = __builtin_fmaf(xv.y, av.y, s0);
          s1 = __builtin_fmaf(xv.y, bw.y, s1);
          s0 = __builtin_fmaf(xv.z, av.z, s0);
          s1 = __builtin_fmaf(xv.z, bw.z, s1);
          s0 = __builtin_fmaf(xv.w, av.w, s0);
          s1 = __builtin_fmaf(xv.w, bw.w, s1);
        }
      } else {
        for (int q = 0; q < D; ++q) {
          s0 = __builtin_fmaf(xr[q], r0[q], s0);
          s1 = __builtin_fmaf(xr[q], r1[q], s1);
        }
      }
      if (!l0) s0 = -INFINITY;
      if (!l1) s1 = -INFINITY;
    }
    const float s2 = __shfl_xor(s0, 32);
    bool fall = false;
    if (ok && h == 0) {
      const float sv[3] = {s0, s1, s2};
      float sb = -INFINITY;
      int lab = -1;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int cj = bi[q];
        if (cj < 0 || cj >= K) continue;
        if (sv[q] > sb || (sv[q] == sb && cj < lab)) { sb = sv[q]; lab = cj; }
      }
      const float xn = xnb[p].x;
      const float xc = xn * cmax;
      const float ub = 3.90625e-3f;  // 2^-8
      const float Df = (float)D;
      const float g32 = Df * 5.9604645e-8f * 1.0001f;                 // gamma_D, u = 2^-24
      const float gmf = (3.0f * Df + 1.0f) * 2.3841858e-7f * 1.001f;  // gamma'_{3D+1}, u = 2^-22
      const float split = ub * ub * (3.0f + 2.0f * ub) * xc;
      const float mfe = gmf * (0.5f * ncmax + (1.0f + ub) * (1.0f + ub) * (1.0f + 2.0f * ub) * xc);
      const float fpe = g32 * (0.5f * ncmax + xc);
      const float E = (split + mfe + fpe) * 1.001f + 1e-30f * Df;
      const bool has4 = bv[3] > -FLT_MAX * 0.5f;
      const float a4 = has4 ? bv[3] : -INFINITY;
      const float slack = (fabsf(bv[0]) + (has4 ? fabsf(a4) : 0.0f)) * 1.6e-5f;
      const float thr = has4 ? a4 + E + slack : -INFINITY;
      if (!tie && lab >= 0 && sb > thr && sb > -INFINITY) {
        labels[p] = lab;
        if (xbo) {
          // drift bounds as in k_assign_mfma: the label's chain minus E bounds
          // its exact score from below; the other candidates' chains plus E
          // and the fourth approximate score plus E and the slack bound the
          // rest from above; the runner-up rr is the other candidate with the
          // larger chain, and hi3 leaves it out
          float hi = a4 + E + slack, hi3 = a4 + E + slack, sr = -INFINITY;
          int rr = -1;
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (bi[q] != lab && bi[q] >= 0 && bi[q] < K) {
              hi = fmaxf(hi, sv[q] + E);
              if (rr < 0 || sv[q] > sr) {
                if (rr >= 0) hi3 = fmaxf(hi3, sr + E);
                rr = bi[q];
                sr = sv[q];
              } else {
                hi3 = fmaxf(hi3, sv[q] + E);
              }
            }
          xbo[p] = make_float4(dist_upper(xn, sb - E), dist_lower(xn, hi, Df),
                               hi3 == -INFINITY ? INFINITY : dist_lower(xn, hi3, Df),
                               __int_as_float(rr));
        }
      } else {
        fall = true;
      }
    }
    const uint64_t mk = __ballot(fall);
    if (mk) {
      const int lead = __ffsll((long long)mk) - 1;
      int base = 0;
      if (lane == lead) base = atomicAdd(nlist2, __popcll(mk));
      base = __shfl(base, lead);
      if (fall) list2[base + __popcll(mk & below)] = (int)p;
    }
    __syncthreads();  // the tile buffers are refilled by the next group
  }
}

// ------------------------------------------------------------------ host
void mfma_free(Engine *e) {
  MfmaState *s = st(e);
  if (!s) return;
  hipFree(s->Xh);
  hipFree(s->Ch);
  hipFree(s->bias);
  hipFree(s->bnd);
  hipFree(s->xnb);
  hipFree(s->dcn);
  hipFree(s->flags);
  hipFree(s->ncand);
  hipFree(s->nflag);
  hipFree(s->flags2);
  hipFree(s->Cl);
  if (s->fp16) {
    hipFree(s->Chb);
    hipFree(s->Clb);
    hipFree(s->biasb);
  }
  hipFree(s->nflag2);
  hipFree(s->kbest);
  hipFree(s->ttab);
  hipFree(s->tdist);
  void *bb[] = {s->xb, s->Cref, s->drift, s->ulist, s->fcnt};
  for (void *b : bb) hipFree(b);
  hipFree(s->home);
  for (hipEvent_t ev : s->kev) hipEventDestroy(ev);
  if (s->pair_stream) {
    hipStreamSynchronize(s->pair_stream);
    hipStreamDestroy(s->pair_stream);
  }
  if (s->pair_fork) hipEventDestroy(s->pair_fork);
  if (s->pair_join) hipEventDestroy(s->pair_join);
  delete s;
  e->mfma = nullptr;
}

void mfma_invalidate(Engine *e) {
  if (st(e)) {
    mfma_pair_join(e);  // the data about to be replaced is still being read
    st(e)->x_ready = false;
    st(e)->bnd_valid = false;
  }
}

void mfma_bounds_invalidate(Engine *e) {
  if (st(e)) st(e)->bnd_valid = false;
}

// centres set from outside the iteration: the drift bounds no longer apply
void mfma_centers_changed(Engine *e, bool new_range) {
  MfmaState *s = st(e);
  if (!s) return;
  s->c_ready = false;
  s->bnd_valid = false;
  if (new_range && s->fp16) s->x_ready = false;  // the FP16 scale may change
}

static hipError_t ensure_state(Engine *e) {
  if (e->mfma) return hipSuccess;
  MfmaState *s = new MfmaState();
  e->mfma = s;
  s->fp16 = e->opt.precision == KM_PREC_FP16;
  s->KS = (e->n + 15) / 16;
  if (s->KS > 8) s->KS = 16;
  else if (s->KS > 6) s->KS = 8;
  else if (s->KS > 4) s->KS = 6;
  else if (s->KS > 2) s->KS = 4;
  s->Dp = s->KS * 16;
  s->Kp = (int)km_cdiv(e->k, MF_KPAD) * MF_KPAD;
  hipError_t r;
  const long xrows = e->streaming ? e->tile : e->m;  // points held in HBM at once
  if ((r = hipMalloc(&s->Xh, sizeof(uint16_t) * (size_t)xrows * s->Dp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->Ch, sizeof(uint16_t) * (size_t)s->Kp * s->Dp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->Cl, sizeof(uint16_t) * (size_t)s->Kp * s->Dp)) != hipSuccess) return r;
  if (s->fp16) {
    if ((r = hipMalloc(&s->Chb, sizeof(uint16_t) * (size_t)s->Kp * s->Dp)) != hipSuccess) return r;
    if ((r = hipMalloc(&s->Clb, sizeof(uint16_t) * (size_t)s->Kp * s->Dp)) != hipSuccess) return r;
    if ((r = hipMalloc(&s->biasb, sizeof(float) * s->Kp)) != hipSuccess) return r;
  } else {
    s->Chb = s->Ch;
    s->Clb = s->Cl;
  }
  if ((r = hipMalloc(&s->bias, sizeof(float) * s->Kp)) != hipSuccess) return r;
  if (!s->fp16) s->biasb = s->bias;
  if ((r = hipMalloc(&s->bnd, sizeof(float) * 4)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->xnb, sizeof(float2) * (size_t)xrows)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->dcn, sizeof(float) * s->Kp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->flags, sizeof(int) * (size_t)xrows)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->flags2, sizeof(int) * (size_t)xrows)) != hipSuccess) return r;
  s->nflag_slots = 64;
  if ((r = hipMalloc(&s->nflag, sizeof(int) * s->nflag_slots)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->ncand, sizeof(int) * NC_SLOT * s->nflag_slots)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->nflag2, sizeof(int) * s->nflag_slots)) != hipSuccess) return r;
  if ((r = hipMemsetAsync(s->ncand, 0, sizeof(int) * NC_SLOT * s->nflag_slots, e->stream)) != hipSuccess)
    return r;
  if ((r = hipMemsetAsync(s->nflag2, 0, sizeof(int) * s->nflag_slots, e->stream)) != hipSuccess)
    return r;
  const size_t kb = sizeof(unsigned long long) * (size_t)split_rows_cap();
  if ((r = hipMalloc(&s->kbest, kb)) != hipSuccess) return r;
  if ((r = hipMemsetAsync(s->kbest, 0, kb, e->stream)) != hipSuccess) return r;
  return hipMemsetAsync(s->nflag, 0, sizeof(int) * s->nflag_slots, e->stream);
}

// Per point upper bounds of |x_i| and of |x_i - xh_i/s| (xh the BF16/FP16
// rounding of s*x_i, exactly as k_convert_rows makes it; s*x - float(xh) is
// exact in FP32).  The FP32 chains of squares are within gamma_d of the exact
// sums (factor 1 + 1.2e-7 d covers it for d < 2^22), one more ulp each for
// the product and the square root.  One point per lane, rows read in 16-byte
// pieces (whole cache lines).
__device__ inline float round_err(float v, int fp16) {
  const uint16_t hi = fp16 ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
  return v - (fp16 ? (float)__builtin_bit_cast(_Float16, hi) : __uint_as_float((uint32_t)hi << 16));
}
__global__ void k_xnorm(long m, int d, const float *__restrict__ X, float scale, int fp16,
                        float2 *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const float *x = X + i * d;
  float s = 0.0f, t = 0.0f;
  if ((d & 3) == 0 && ((uintptr_t)X & 15) == 0) {
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
#pragma unroll 8
    for (int q = 0; q < d / 4; ++q) {
      const float4 v = x4[q];
      s = __builtin_fmaf(v.x, v.x, s);
      s = __builtin_fmaf(v.y, v.y, s);
      s = __builtin_fmaf(v.z, v.z, s);
      s = __builtin_fmaf(v.w, v.w, s);
      const float r0 = round_err(v.x * scale, fp16), r1 = round_err(v.y * scale, fp16);
      const float r2 = round_err(v.z * scale, fp16), r3 = round_err(v.w * scale, fp16);
      t = __builtin_fmaf(r0, r0, t);
      t = __builtin_fmaf(r1, r1, t);
      t = __builtin_fmaf(r2, r2, t);
      t = __builtin_fmaf(r3, r3, t);
    }
  } else {
    for (int q = 0; q < d; ++q) {
      s = __builtin_fmaf(x[q], x[q], s);
      const float r = round_err(x[q] * scale, fp16);
      t = __builtin_fmaf(r, r, t);
    }
  }
  const float inf = 1.0f + (float)d * 1.2e-7f + 1e-6f;
  out[i] = make_float2(sqrtf(s * inf) * (1.0f + 1e-6f) + 1e-30f,
                       sqrtf(t * inf) * (1.0f + 1e-6f) / scale + 1e-30f);
}

// max |v| over a device array (FP16 scaling only; one host read per run)
__global__ void k_absmax(long n, const float *__restrict__ v, unsigned *__restrict__ out) {
  float m = 0.0f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(v[i]));
  atomicMax(out, __float_as_uint(m));
}

static hipError_t prepare_centres(Engine *e);

static hipError_t prepare_x(Engine *e) {
  MfmaState *s = st(e);
  if (s->x_ready) return hipSuccess;
  hipError_t r;
  s->scale = 1.0f;
  if (s->fp16) {
    unsigned *dm = nullptr;
    if ((r = hipMalloc(&dm, sizeof(unsigned))) != hipSuccess) return r;
    const unsigned hx = __builtin_bit_cast(unsigned, e->x_absmax);  // streamed data: host pass
    hipMemcpyAsync(dm, &hx, sizeof(unsigned), hipMemcpyHostToDevice, e->stream);
    if (!e->streaming)
      hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, e->stream, (long)e->m * e->n, e->dX, dm);
    hipLaunchKernelGGL(k_absmax, dim3(64), dim3(256), 0, e->stream, (long)e->k * e->n, e->dC, dm);
    unsigned um = 0;
    hipMemcpyAsync(&um, dm, sizeof(unsigned), hipMemcpyDeviceToHost, e->stream);
    if ((r = hipStreamSynchronize(e->stream)) != hipSuccess) return r;
    hipFree(dm);
    float mx = __builtin_bit_cast(float, um);
    // largest power of two with mx * scale <= 2^14 (FP16 max is 65504)
    int ex = 0;
    if (mx > 0.0f && isfinite(mx)) frexpf(mx, &ex);
    s->scale = ldexpf(1.0f, 14 - ex);
  }
  if (!e->streaming) {
    const long tot = (long)e->m * s->Dp;
    hipLaunchKernelGGL(k_convert_rows, dim3(km_grid(tot, 256)), dim3(256), 0, e->stream, e->m,
                       e->n, s->Dp, e->dX, s->scale, (int)s->fp16, s->Xh);
    hipLaunchKernelGGL(k_xnorm, dim3(km_cdiv(e->m, 256)), dim3(256), 0, e->stream, e->m, e->n,
                       e->dX, s->scale, (int)s->fp16, s->xnb);
    if ((r = hipGetLastError()) != hipSuccess) return r;
  }
  s->x_ready = true;
  return hipSuccess;
}

// Streamed tiles: reduced-precision copy and norm bounds of rows [0, rows)
// of a tile buffer into the state's tile-sized arrays.
hipError_t mfma_prepare_rows(Engine *e, const float *X, long rows) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  if (!s->x_ready)
    if ((r = prepare_x(e)) != hipSuccess) return r;
  if (!s->c_ready)
    if ((r = prepare_centres(e)) != hipSuccess) return r;
  if (rows <= 0) return hipSuccess;
  const long tot = rows * s->Dp;
  hipLaunchKernelGGL(k_convert_rows, dim3(km_grid(tot, 256)), dim3(256), 0, e->stream, rows, e->n,
                     s->Dp, X, s->scale, (int)s->fp16, s->Xh);
  hipLaunchKernelGGL(k_xnorm, dim3(km_cdiv(rows, 256)), dim3(256), 0, e->stream, rows, e->n, X,
                     s->scale, (int)s->fp16, s->xnb);
  return hipGetLastError();
}

static hipError_t prepare_centres(Engine *e) {
  MfmaState *s = st(e);
  const unsigned pg = (unsigned)std::min<long>(km_cdiv(s->Kp, 4), 1024);  // a wave per row
  hipLaunchKernelGGL(k_prepare_centres, dim3(pg), dim3(256), 0, e->stream, e->k, e->n, s->Dp, s->Kp,
                     e->dC, e->dnc, s->scale, (int)s->fp16, s->Ch, s->Cl, s->bias, s->dcn, s->Chb,
                     s->Clb, s->biasb);
  hipLaunchKernelGGL(k_centre_bounds, dim3(1), dim3(256), 0, e->stream, e->k, e->n, e->dnc, s->dcn,
                     s->bnd);
  s->c_ready = true;
  // tile table of the label-ordered assign, which runs when there are no
  // drift bounds to use (after centres set from outside)
  s->ttab_ok = false;
  if (e->dorder && e->k <= MF_ORD_MAX_K && !s->bnd_valid) {
    hipError_t r;
    const int nKT = s->Kp / MF_TILE;
    if (!s->ttab &&
        (r = hipMalloc(&s->ttab, sizeof(int) * (size_t)e->k * nKT)) != hipSuccess)
      return r;
    if (!s->tdist &&
        (r = hipMalloc(&s->tdist, sizeof(float) * (size_t)e->k * nKT)) != hipSuccess)
      return r;
    hipLaunchKernelGGL(k_tile_order, dim3(e->k), dim3(256), sizeof(float) * (e->n + nKT),
                       e->stream, e->k, e->n, nKT, e->dCt, e->ldc, e->dnc, s->ttab, s->tdist);
    s->ttab_ok = true;
  }
  return hipGetLastError();
}

// called after every centre update: centres are means of points, so their
// magnitude stays within the range the FP16 scale was chosen for
hipError_t mfma_centers(Engine *e) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  if (!st(e)->x_ready) {
    st(e)->c_ready = false;
    return hipSuccess;
  }
  return prepare_centres(e);
}

// order/ttab/home all non-null: the label-ordered variant; order alone: the
// points order[0..M) (with Mdev, the first *Mdev of them) in that order
template <int KS, bool FP16, bool ORD, bool LIST>
static void launch_mfma_mode(hipStream_t strm, dim3 g, size_t lds, long M, const uint16_t *Xh,
                             const float *X, int D, int K, int Kp, const uint16_t *Ch,
                             const float *bias, const float *C, const float *nc, const float *bnd,
                             const float2 *xnb, float inv_s2, float eta, int *labels, int *flags,
                             int *nflag, int *ncand, const int *stop, const int *order,
                             const int *ttab, const float *tdist, const int *home, const int *Mdev,
                             float4 *xbo) {
  static int attr_lds = -1;
  if ((int)lds > attr_lds) {
    hipFuncSetAttribute((const void *)k_assign_mfma<KS, FP16, ORD, LIST>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_lds = (int)lds;
  }
  hipLaunchKernelGGL((k_assign_mfma<KS, FP16, ORD, LIST>), g, dim3(64 * MF_WAVES), lds, strm, M,
                     Xh, X, D, K, Kp, Ch, bias, C, nc, bnd, xnb, inv_s2, eta, FLT_MAX, labels,
                     flags, nflag, ncand, stop, order, ttab, tdist, home, Mdev, xbo);
}

template <int KS>
static void launch_mfma_ks(hipStream_t strm, bool fp16, dim3 g, long M, const uint16_t *Xh,
                           const float *X, int D, int K, int Kp, const uint16_t *Ch,
                           const float *bias, const float *C, const float *nc, const float *bnd,
                           const float2 *xnb, float inv_s2, float eta, int *labels,
                           int *flags, int *nflag, int *ncand, const int *stop, const int *order,
                           const int *ttab, const float *tdist, const int *home, const int *Mdev,
                           float4 *xbo) {
  const bool ord = order != nullptr && home != nullptr;
  const bool list = order != nullptr && home == nullptr;
  const size_t lds = 2 * (size_t)MF_STAGE * MF_TILE * (KS * 2) * 16 + 2 * MF_STAGE * MF_TILE * 4 +
                     (ord ? 2 * sizeof(int) * (size_t)(Kp / MF_TILE) : 0);
#define KM_ARGS                                                                                 \
  strm, g, lds, M, Xh, X, D, K, Kp, Ch, bias, C, nc, bnd, xnb, inv_s2, eta, labels, flags, nflag, \
      ncand, stop, order, ttab, tdist, home, Mdev, xbo
  if (fp16) {
    if (ord) launch_mfma_mode<KS, true, true, false>(KM_ARGS);
    else if (list) launch_mfma_mode<KS, true, false, true>(KM_ARGS);
    else launch_mfma_mode<KS, true, false, false>(KM_ARGS);
  } else {
    if (ord) launch_mfma_mode<KS, false, true, false>(KM_ARGS);
    else if (list) launch_mfma_mode<KS, false, false, true>(KM_ARGS);
    else launch_mfma_mode<KS, false, false, false>(KM_ARGS);
  }
#undef KM_ARGS
}

// Drift-bound buffers (once per engine).
static hipError_t ensure_bounds(Engine *e) {
  MfmaState *s = st(e);
  if (s->xb) return hipSuccess;
  hipError_t r;
  if ((r = hipMalloc(&s->xb, sizeof(float4) * (size_t)e->m)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->Cref, sizeof(float) * (size_t)e->k * e->n)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->drift, sizeof(float) * (size_t)(e->k + 1))) != hipSuccess) return r;
  if ((r = hipMemsetAsync(s->drift, 0, sizeof(float) * (size_t)(e->k + 1), e->stream)) != hipSuccess)
    return r;
  if ((r = hipMalloc(&s->ulist, sizeof(int) * (size_t)e->m)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->fcnt, sizeof(int) * 2)) != hipSuccess) return r;
  if ((r = hipStreamCreateWithFlags(&s->pair_stream, hipStreamNonBlocking)) != hipSuccess) return r;
  if ((r = hipEventCreateWithFlags(&s->pair_fork, hipEventDisableTiming)) != hipSuccess) return r;
  if ((r = hipEventCreateWithFlags(&s->pair_join, hipEventDisableTiming)) != hipSuccess) return r;
  // lists of zero points until the first filter (a launch whose filter a
  // convergence stop skipped reads these)
  return hipMemsetAsync(s->fcnt, 0, sizeof(int) * 2, e->stream);
}

// After an assign over all resident points every point's bounds hold for the
// current centres (fresh ones for the assigned points, moved ones for the
// others): they become the reference centres.
static hipError_t ref_copy(Engine *e, MfmaState *s, long rows) {
  (void)rows;
  const long nc = (long)e->k * e->n;
  hipLaunchKernelGGL(k_ref_copy, dim3(km_cdiv(nc, 256)), dim3(256), 0, e->stream, nc, e->dC,
                     s->Cref, s->drift + e->k, e->assign_stop);
  s->bnd_valid = true;
  return hipGetLastError();
}

hipError_t mfma_assign(Engine *e, long r0, long rows) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  // the previous call's pair decisions (if nothing joined them yet) read the
  // buffers prepared and counted below
  if ((r = mfma_pair_join(e)) != hipSuccess) return r;
  if (e->n > 256) return hipErrorNotSupported;
  if (!s->x_ready) {
    if ((r = prepare_x(e)) != hipSuccess) return r;
    s->c_ready = false;
  }
  if (!s->c_ready)
    if ((r = prepare_centres(e)) != hipSuccess) return r;
  const bool all = r0 == 0 && rows == e->m;
  if (all && (r = ensure_bounds(e)) != hipSuccess) return r;
  // this iteration's counters start at zero before the filter adds to them
  if ((r = mfma_ensure_slots(e, e->cur_slot + 1)) != hipSuccess) return r;
  int *nc2 = s->ncand + (size_t)e->cur_slot * NC_SLOT;
  if ((r = hipMemsetAsync(s->nflag + e->cur_slot, 0, sizeof(int), e->stream)) != hipSuccess)
    return r;
  if ((r = hipMemsetAsync(nc2, 0, sizeof(int) * NC_SLOT, e->stream)) != hipSuccess) return r;
  const int *order = nullptr, *home = nullptr, *Mdev = nullptr;
  bool pairs = false;  // two-candidate points to decide after the assign
  if (all && s->bnd_valid) {
    // the drift bounds of the last assign over all points: only the points
    // they leave open are assigned, in the order of the filter's runs (the
    // filter reads every point's label and bounds once, coalesced, and lists
    // the points with two candidates for k_pair_decide)
    hipLaunchKernelGGL(k_drift, dim3(km_cdiv(e->k, 4)), dim3(256), 0, e->stream, e->k, e->n,
                       e->dC, s->Cref, s->drift, s->fcnt, e->assign_stop);
    hipLaunchKernelGGL(k_bound_filter, dim3(km_cdiv(rows, 256 * BF_PPT)), dim3(256), 0, e->stream,
                       rows, (const int *)nullptr, e->dlabels, s->xb, s->drift, e->k, s->bnd,
                       s->xnb, e->n, rows, s->ulist, s->fcnt, e->assign_stop);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    pairs = true;
    order = s->ulist;
    Mdev = s->fcnt;
  } else if (all && e->order_ok && s->ttab_ok && e->k <= MF_ORD_MAX_K) {
    // all resident points, with the previous update's order and a tile
    // table: the label-ordered variant
    const long nwg = km_cdiv(rows, MF_PTS);
    if (nwg > s->home_cap) {
      hipFree(s->home);
      s->home = nullptr;
      s->home_cap = 0;
      if ((r = hipMalloc(&s->home, sizeof(int) * (size_t)nwg)) != hipSuccess) return r;
      s->home_cap = nwg;
    }
    hipLaunchKernelGGL(k_wg_home, dim3(km_cdiv(nwg, 256)), dim3(256), 0, e->stream, nwg, rows,
                       (const int *)nullptr, e->k, e->dorder, e->dlabels, s->home);
    order = e->dorder;
    home = s->home;
  }
  if ((r = mfma_assign_view(e, e->dX + r0 * e->n, s->Xh + r0 * s->Dp, s->xnb + r0,
                            e->dlabels + r0, rows, order, home, Mdev, all ? s->xb : nullptr,
                            false)) != hipSuccess)
    return r;
  if (pairs) {
    // the two-candidate points (disjoint from the open ones the assign
    // scored and the re-resolve revisits; the counters are atomics) are
    // decided on the side stream while the re-resolve runs; mfma_reresolve
    // (or mfma_pair_join) makes the engine's stream wait for them
    if ((r = hipEventRecord(s->pair_fork, e->stream)) != hipSuccess) return r;
    if ((r = hipStreamWaitEvent(s->pair_stream, s->pair_fork, 0)) != hipSuccess) return r;
    const unsigned pg = (unsigned)std::min<long>(km_cdiv(rows, 256), (long)num_cus() * 16);
    hipLaunchKernelGGL(k_pair_decide, dim3(pg), dim3(256), 0, s->pair_stream, s->ulist,
                       s->fcnt + 1, rows, e->dlabels, s->xb, s->xnb, s->bnd, e->n,
                       (const float *)e->dX, (const float *)e->dC, (const float *)e->dnc, nc2,
                       e->assign_stop);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if ((r = hipEventRecord(s->pair_join, s->pair_stream)) != hipSuccess) return r;
    s->pair_pending = true;
  }
  if (!all) {
    s->bnd_valid = false;  // a partial assign leaves the bounds unusable
    return hipSuccess;
  }
  return ref_copy(e, s, rows);
}

// Timing runs: an event before the assign kernel(s) of one call, and the
// matching one after them, on the stream they run on (collected by
// collect_kernel_times).
static hipError_t timing_begin(Engine *e, MfmaState *s, hipEvent_t *k1) {
  *k1 = nullptr;
  if (!e->opt.timing) return hipSuccess;
  if (s->kev_used >= 4096) collect_kernel_times(e, s);  // step calls outside a run
  hipError_t r;
  while (s->kev.size() < s->kev_used + 2) {
    hipEvent_t ev = nullptr;
    if ((r = hipEventCreate(&ev)) != hipSuccess) return r;
    s->kev.push_back(ev);
  }
  *k1 = s->kev[s->kev_used + 1];
  r = hipEventRecord(s->kev[s->kev_used], e->stream);
  s->kev_used += 2;
  return r;
}
static hipError_t timing_end(Engine *e, hipEvent_t k1) {
  return k1 ? hipEventRecord(k1, e->stream) : hipSuccess;
}

// One launch of the assign kernel (see mfma_assign_view).
static hipError_t launch_assign(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                                int *lab, long rows, const int *order, const int *home,
                                const int *Mdev, float4 *xbo) {
  MfmaState *s = st(e);
  int *nf = s->nflag + e->cur_slot;
  int *nc2 = s->ncand + (size_t)e->cur_slot * NC_SLOT;
  const float eta = s->fp16 ? 2.9802322e-8f / s->scale : 0.0f;  // half FP16 subnormal step
  const float inv_s2 = 1.0f / (s->scale * s->scale);
  dim3 g(km_cdiv(rows, MF_PTS));
  const int *tt = home ? s->ttab : nullptr;
  const float *td = home ? s->tdist : nullptr;
#define KM_LAUNCH(KS)                                                                              \
  launch_mfma_ks<KS>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, \
                     e->dnc, s->bnd, xn, inv_s2, eta, lab, s->flags, nf, nc2, e->assign_stop,      \
                     order, tt, td, home, Mdev, xbo)
  switch (s->KS) {
    case 1: KM_LAUNCH(1); break;
    case 2: KM_LAUNCH(2); break;
    case 4: KM_LAUNCH(4); break;
    case 6: KM_LAUNCH(6); break;
    case 8: KM_LAUNCH(8); break;
    default: KM_LAUNCH(16); break;
  }
#undef KM_LAUNCH
  return hipGetLastError();
}

// Assignment of `rows` points: FP32 rows X, their reduced-precision copy Xh
// and norm bounds xn, labels written to lab (all indexed from 0).  order and
// home (nullable, both or neither): the label-ordered variant over all rows;
// Mdev (nullable, with order): the number of positions of order to assign;
// xbo (nullable): fresh drift bounds of the assigned points.
hipError_t mfma_assign_view(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                            int *lab, long rows, const int *order, const int *home,
                            const int *Mdev, float4 *xbo, bool zero_counters) {
  hipError_t r;
  MfmaState *s = st(e);
  if (!s->c_ready)
    if ((r = prepare_centres(e)) != hipSuccess) return r;
  if ((r = mfma_ensure_slots(e, e->cur_slot + 1)) != hipSuccess) return r;
  if (zero_counters) {
    if ((r = hipMemsetAsync(s->nflag + e->cur_slot, 0, sizeof(int), e->stream)) != hipSuccess)
      return r;
    if ((r = hipMemsetAsync(s->ncand + (size_t)e->cur_slot * NC_SLOT, 0, sizeof(int) * NC_SLOT,
                            e->stream)) != hipSuccess)
      return r;
  }
  if (rows <= 0) return hipSuccess;
  hipEvent_t k1;
  if ((r = timing_begin(e, s, &k1)) != hipSuccess) return r;
  if ((r = launch_assign(e, X, Xh, xn, lab, rows, order, home, Mdev, xbo)) != hipSuccess)
    return r;
  return timing_end(e, k1);
}

hipError_t mfma_ensure_slots(Engine *e, int n) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  if (n <= s->nflag_slots) return hipSuccess;
  int *arrs[3] = {s->nflag, s->ncand, s->nflag2};
  const size_t per[3] = {1, NC_SLOT, 1};  // ints per slot
  for (int a = 0; a < 3; ++a) {
    int *nb = nullptr;
    if ((r = hipMalloc(&nb, sizeof(int) * per[a] * n)) != hipSuccess) return r;
    if ((r = hipMemsetAsync(nb, 0, sizeof(int) * per[a] * n, e->stream)) != hipSuccess) return r;
    if ((r = hipMemcpyAsync(nb, arrs[a], sizeof(int) * per[a] * s->nflag_slots, hipMemcpyDeviceToDevice,
                            e->stream)) != hipSuccess)
      return r;
    if ((r = hipStreamSynchronize(e->stream)) != hipSuccess) return r;
    hipFree(arrs[a]);
    arrs[a] = nb;
  }
  s->nflag = arrs[0];
  s->ncand = arrs[1];
  s->nflag2 = arrs[2];
  s->nflag_slots = n;
  return hipSuccess;
}

// One-time data preparation (reduced-precision copy of X, row norms) so the
// iteration loop does no per-dataset work; called when data and centres are set.
hipError_t mfma_prepare(Engine *e) {
  hipError_t r;
  if (!e->dX || !e->centers_set || e->n > 256) return hipSuccess;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  if (!s->x_ready) {
    if ((r = prepare_x(e)) != hipSuccess) return r;
    s->c_ready = false;
  }
  if (!s->c_ready)
    if ((r = prepare_centres(e)) != hipSuccess) return r;
  return hipStreamSynchronize(e->stream);
}

// FP32 full search for the uncertified points (device-side count).
hipError_t mfma_reresolve(Engine *e, long r0, long rows) {
  MfmaState *s = st(e);
  const bool all = r0 == 0 && rows == e->m;
  const hipError_t r = mfma_reresolve_view(e, e->dX + r0 * e->n, s->Xh + r0 * s->Dp, s->xnb + r0,
                                           e->dlabels + r0, rows, all ? s->xb : nullptr);
  const hipError_t rj = mfma_pair_join(e);  // whatever the re-resolve's status
  return r != hipSuccess ? r : rj;
}

hipError_t mfma_pair_join(Engine *e) {
  MfmaState *s = st(e);
  if (!s || !s->pair_pending) return hipSuccess;
  s->pair_pending = false;
  return hipStreamWaitEvent(e->stream, s->pair_join, 0);
}

// Points left open by the assign kernel: deep MFMA pass, then the FP32 full
// search of what remains (device-side counts; X, Xh, xn are the same rows
// the assign pass saw, indexed from 0).
hipError_t mfma_reresolve_view(Engine *e, const float *X, const uint16_t *Xh, const float2 *xn,
                               int *lab, long rows, float4 *xbo) {
  MfmaState *s = st(e);
  hipError_t r;
  if ((r = hipMemsetAsync(s->nflag2 + e->cur_slot, 0, sizeof(int), e->stream)) != hipSuccess) return r;
  if (rows <= 0) return hipSuccess;
  (void)Xh;  // the deep pass splits the FP32 rows itself
  const long g = std::min<long>(km_cdiv(rows, 128), (long)num_cus() * 8);
#define DEEP(KS)                                                                                 \
  hipLaunchKernelGGL(k_deep_mfma<KS>, dim3(g), dim3(256), 0, e->stream, s->nflag + e->cur_slot,  \
                     s->flags, X, e->n, e->k, s->Kp, s->Chb, s->Clb, s->biasb, e->dC, e->dnc,     \
                     s->bnd, xn, FLT_MAX, lab, s->flags2, s->nflag2 + e->cur_slot, xbo)
  switch (s->KS) {
    case 1: DEEP(1); break;
    case 2: DEEP(2); break;
    case 4: DEEP(4); break;
    case 6: DEEP(6); break;
    case 8: DEEP(8); break;
    default: DEEP(16); break;
  }
#undef DEEP
  if ((r = hipGetLastError()) != hipSuccess) return r;
  return launch_assign_f32(e->stream, rows, s->nflag2 + e->cur_slot, e->n, X, s->flags2, e->k,
                           e->dCt, e->ldc, e->dnc, lab, nullptr, nullptr, true, s->kbest);
}

const uint16_t *mfma_tile_xh(Engine *e) { return st(e)->Xh; }
const float2 *mfma_tile_xnorm(Engine *e) { return st(e)->xnb; }
int mfma_row_stride(Engine *e) { return st(e)->Dp; }
bool mfma_view(Engine *e, MfmaView *v) {
  MfmaState *s = st(e);
  if (!s || !s->x_ready || !s->c_ready || e->streaming) return false;
  *v = MfmaView{s->Xh, s->Ch, s->bias, s->bnd, s->xnb, s->Dp, s->KS, s->Kp, s->fp16};
  return true;
}
bool mfma_ready(Engine *e) { return st(e) && st(e)->x_ready && !e->streaming; }

// Assign kernel durations recorded so far (timing) into the stats.
static void collect_kernel_times(Engine *e, MfmaState *s) {
  if (!s->kev_used) return;
  if (hipStreamSynchronize(e->stream) == hipSuccess)
    for (size_t i = 0; i + 1 < s->kev_used; i += 2) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, s->kev[i], s->kev[i + 1]) == hipSuccess) {
        e->stats.ms_assign_kernel += ms;
        e->stats.assign_launches += 1;
      }
    }
  s->kev_used = 0;
}

void mfma_collect(Engine *e, int nslots) {
  MfmaState *s = st(e);
  if (!s) return;
  collect_kernel_times(e, s);
  if (nslots <= 0) return;
  std::vector<int> v(nslots);
  const int n = nslots < s->nflag_slots ? nslots : s->nflag_slots;
  if (hipMemcpyAsync(v.data(), s->nflag2, sizeof(int) * n, hipMemcpyDeviceToHost, e->stream) !=
      hipSuccess)
    return;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return;
  for (int i = 0; i < n; ++i) {
    e->stats.reresolved += v[i];
    e->last_reresolved = v[i];
  }
  if (hipMemcpyAsync(v.data(), s->nflag, sizeof(int) * n, hipMemcpyDeviceToHost, e->stream) !=
      hipSuccess)
    return;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return;
  for (int i = 0; i < n; ++i) e->stats.deep += v[i];
  std::vector<int> w((size_t)n * NC_SLOT);
  if (hipMemcpyAsync(w.data(), s->ncand, sizeof(int) * w.size(), hipMemcpyDeviceToHost, e->stream) !=
      hipSuccess)
    return;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return;
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < NC_SPREAD; ++c) {
      e->stats.rescored += w[(size_t)i * NC_SLOT + c * NC_STRIDE];
      e->stats.tiles += w[(size_t)i * NC_SLOT + c * NC_STRIDE + 1];
      // slot 3: the points the filter did not leave open; slot 4: those of
      // them it decided between two candidates
      e->stats.bounded += w[(size_t)i * NC_SLOT + c * NC_STRIDE + 3] -
                          w[(size_t)i * NC_SLOT + c * NC_STRIDE + 4];
      e->stats.paired += w[(size_t)i * NC_SLOT + c * NC_STRIDE + 4];
      e->stats.scored += w[(size_t)i * NC_SLOT + c * NC_STRIDE + 5];
      e->stats.tiles_total += w[(size_t)i * NC_SLOT + c * NC_STRIDE + 2];
    }
}

}  // namespace km
