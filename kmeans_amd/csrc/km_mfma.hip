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
// ed., sec. 3.1; u = 2^-24, gamma_n = n u / (1 - n u)):
//   |fl-chain - exact| <= gamma_D (0.5 nc_j + sum|x c|)          (FP32 path)
//   |xh ch - s^2 x.c|  <= (2 ub + ub^2) s^2 sum|x c| + eta terms  (rounding)
//   |mfma - xh.ch|     <= gamma'_{D+1} (0.5 s^2 nc_j + sum|xh ch|) with a
//                         pessimistic unit roundoff 2^-22 for the matrix core
// and sum|x c| <= |x| |c| (Cauchy-Schwarz), maximised over j with the
// largest centre norm.  ub = 2^-8 (BF16) or 2^-11 (FP16).
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
#define MF_PTS 256     // points per workgroup (4 waves x 2 groups x 32)

// Round-to-nearest-even conversions written out so the rounding is known.
__device__ inline uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline uint16_t f32_to_f16_rne(float f) {
  _Float16 h = (_Float16)f;  // IEEE conversion, round to nearest even
  return __builtin_bit_cast(uint16_t, h);
}

struct MfmaState {
  int Dp = 0;        // padded dimension (multiple of 16)
  int KS = 0;        // k-steps of 16
  int Kp = 0;        // centres padded to MF_TILE
  bool fp16 = false;
  float scale = 1.f; // power of two applied to X and C before conversion (FP16)
  uint16_t *Xh = nullptr;   // m x Dp
  uint16_t *Ch = nullptr;   // Kp x Dp
  float *bias = nullptr;    // Kp: -0.5*scale^2*nc (or -FLT_MAX past k)
  float *bnd = nullptr;     // [cmax, ncmax]
  int *flags = nullptr;     // m
  int *nflag = nullptr;     // per iteration slot
  int nflag_slots = 0;
  bool x_ready = false;
  bool c_ready = false;
};

static MfmaState *st(Engine *e) { return (MfmaState *)e->mfma; }

// ------------------------------------------------------------------ kernels
__global__ void k_convert_rows(long m, int d, int Dp, const float *__restrict__ X, float scale,
                               int fp16, uint16_t *__restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * Dp) return;
  const long r = e / Dp;
  const int q = (int)(e - r * Dp);
  const float v = q < d ? X[r * d + q] * scale : 0.0f;
  out[e] = fp16 ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
}

// Ch, bias and the bound constants [max |c_j|, max nc_j] (one block).
__global__ __launch_bounds__(256) void k_prepare_centres(int k, int d, int Dp, int Kp,
                                                         const float *__restrict__ C,
                                                         const float *__restrict__ nc, float scale,
                                                         int fp16, uint16_t *__restrict__ Ch,
                                                         float *__restrict__ bias,
                                                         float *__restrict__ bnd) {
  __shared__ float red[256];
  for (long e = threadIdx.x; e < (long)Kp * Dp; e += blockDim.x) {
    const long j = e / Dp;
    const int q = (int)(e - j * Dp);
    const float v = (j < k && q < d) ? C[j * d + q] * scale : 0.0f;
    Ch[e] = fp16 ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
  }
  float mx = 0.0f;
  for (int j = threadIdx.x; j < Kp; j += blockDim.x) {
    bias[j] = j < k ? -0.5f * nc[j] * (scale * scale) : -FLT_MAX;
    if (j < k) mx = fmaxf(mx, nc[j]);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // nc_j is the FP32 chain; the true |c_j|^2 is at most nc_j (1 + gamma_D)
    const float ncmax = red[0] * (1.0f + 1e-5f) + 1e-30f;
    bnd[0] = sqrtf(ncmax) * (1.0f + 1e-6f);
    bnd[1] = ncmax;
  }
}

// top-3 of packed values: b1 >= b2 >= b3
__device__ inline void top3_insert(float p, float &b1, float &b2, float &b3) {
  const float n3 = __builtin_amdgcn_fmed3f(b2, b3, p);
  const float n2 = __builtin_amdgcn_fmed3f(b1, b2, p);
  b1 = fmaxf(b1, p);
  b2 = n2;
  b3 = n3;
}

// row of output register r (0..15) of a 32x32 MFMA tile for lane half h
__device__ inline int mf_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

template <int KS, bool FP16>
__global__ __launch_bounds__(256) void k_assign_mfma(
    long M, const uint16_t *__restrict__ Xh, const float *__restrict__ X, int D, int K, int Kp,
    const uint16_t *__restrict__ Ch, const float *__restrict__ bias, const float *__restrict__ C,
    const float *__restrict__ nc, const float *__restrict__ bnd, float inv_s2, float ub,
    float eta, int *__restrict__ labels, int *__restrict__ flags, int *__restrict__ nflag) {
  constexpr int Dp = KS * 16;
  constexpr int LDA = Dp + 8;  // bf16 elements per LDS row (16-byte pad)
  __shared__ __attribute__((aligned(16))) uint16_t As[2][MF_TILE * LDA];
  __shared__ __attribute__((aligned(16))) float Bs[2][MF_TILE];
  typedef typename std::conditional<FP16, f16x8_t, bf16x8_t>::type vec8;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, col = lane & 31;
  const long p0 = (long)blockIdx.x * MF_PTS + w * 64;
  const int nKT = Kp / MF_TILE;

  // point fragments: lane holds point (p0 + g*32 + col), k = ks*16 + 8h + 0..7
  u32x4_t bfrag[2][KS];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const long p = p0 + g * 32 + col;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (p < M)
        bfrag[g][ks] = *reinterpret_cast<const u32x4_t *>(Xh + p * Dp + ks * 16 + 8 * h);
      else
        bfrag[g][ks] = u32x4_t{0u, 0u, 0u, 0u};
    }
  }

  // centre tiles: global -> registers (issued before the MFMAs of the
  // current tile) -> LDS (after them), double-buffered
  constexpr int CHUNKS = MF_TILE * Dp / 8;  // 16-byte chunks per tile
  constexpr int PER = (CHUNKS + 255) / 256;
  u32x4_t pre[PER];
  float prebias = 0.0f;
  auto fetch_tile = [&](int kt) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int c = tid + 256 * i;
      if (c < CHUNKS) {
        const int r = c / (Dp / 8), q = c - r * (Dp / 8);
        pre[i] = *reinterpret_cast<const u32x4_t *>(Ch + ((long)kt * MF_TILE + r) * Dp + q * 8);
      }
    }
    if (tid < MF_TILE) prebias = bias[kt * MF_TILE + tid];
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int c = tid + 256 * i;
      if (c < CHUNKS) {
        const int r = c / (Dp / 8), q = c - r * (Dp / 8);
        *reinterpret_cast<u32x4_t *>(&As[buf][r * LDA + q * 8]) = pre[i];
      }
    }
    if (tid < MF_TILE) Bs[buf][tid] = prebias;
  };

  float b1[2], b2[2], b3[2];
  int t1[2], t2[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    b1[g] = b2[g] = b3[g] = -FLT_MAX;
    t1[g] = t2[g] = 0;
  }

  fetch_tile(0);
  store_tile(0);
  __syncthreads();
  for (int kt = 0; kt < nKT; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nKT) fetch_tile(kt + 1);
    // bias as the accumulator input: rows 8*(r/4) + 4h + r%4
    f32x16_t cinit;
#pragma unroll
    for (int r = 0; r < 16; ++r) cinit[r] = Bs[buf][mf_row(r, h)];
    f32x16_t acc0 = cinit, acc1 = cinit;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const u32x4_t a = *reinterpret_cast<const u32x4_t *>(&As[buf][col * LDA + ks * 16 + 8 * h]);
      if constexpr (FP16) {
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vec8, a),
                                                      __builtin_bit_cast(vec8, bfrag[0][ks]), acc0,
                                                      0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vec8, a),
                                                      __builtin_bit_cast(vec8, bfrag[1][ks]), acc1,
                                                      0, 0, 0);
      } else {
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vec8, a),
                                                       __builtin_bit_cast(vec8, bfrag[0][ks]),
                                                       acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vec8, a),
                                                       __builtin_bit_cast(vec8, bfrag[1][ks]),
                                                       acc1, 0, 0, 0);
      }
    }
    // epilogue: pack the register id into the low 4 bits, keep the top 3
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const f32x16_t &acc = g == 0 ? acc0 : acc1;
      const float o1 = b1[g], o2 = b2[g];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float p = __uint_as_float((__float_as_uint(acc[r]) & 0xFFFFFFF0u) | (unsigned)r);
        top3_insert(p, b1[g], b2[g], b3[g]);
      }
      const unsigned u1 = __float_as_uint(b1[g]), u2 = __float_as_uint(b2[g]);
      const int nt2 = (u2 == __float_as_uint(o1)) ? t1[g] : (u2 == __float_as_uint(o2)) ? t2[g] : kt;
      t1[g] = (u1 != __float_as_uint(o1)) ? kt : t1[g];
      t2[g] = nt2;
    }
    if (kt + 1 < nKT) store_tile(buf ^ 1);
    __syncthreads();
  }

  // merge the two lane halves (same points, different centre rows)
  const float cmax = bnd[0], ncmax = bnd[1];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    // candidate indices of this half
    const unsigned u1 = __float_as_uint(b1[g]), u2 = __float_as_uint(b2[g]);
    int i1 = t1[g] * MF_TILE + mf_row((int)(u1 & 15u), h);
    int i2 = t2[g] * MF_TILE + mf_row((int)(u2 & 15u), h);
    bool tie = (u1 == u2) || (u2 == __float_as_uint(b3[g]));
    float v1 = b1[g], v2 = b2[g], v3 = b3[g];
    // the other half's triple
    const float w1 = __shfl_xor(v1, 32), w2 = __shfl_xor(v2, 32), w3 = __shfl_xor(v3, 32);
    const int j1 = __shfl_xor(i1, 32), j2 = __shfl_xor(i2, 32);
    const bool otie = __shfl_xor((int)tie, 32) != 0;
    // merge two sorted triples (exact bitwise duplicates are flagged)
    float m1, m2, m3;
    int k1, k2;
    {
      // take in order; both halves are sorted descending
      float a[3] = {v1, v2, v3}, b[3] = {w1, w2, w3};
      int ai[2] = {i1, i2}, bi[2] = {j1, j2};
      int ia = 0, ib = 0;
      float out[3];
      int oi[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const bool takeA = a[ia] >= b[ib];
        out[t] = takeA ? a[ia] : b[ib];
        oi[t] = takeA ? (ia < 2 ? ai[ia] : -1) : (ib < 2 ? bi[ib] : -1);
        if (takeA) ++ia; else ++ib;
      }
      m1 = out[0]; m2 = out[1]; m3 = out[2];
      k1 = oi[0]; k2 = oi[1];
      tie = tie || otie || (__float_as_uint(m1) == __float_as_uint(m2)) ||
            (__float_as_uint(m2) == __float_as_uint(m3));
      // a candidate index of -1 can only appear if a half's third entry
      // outranked the other half's first two, which then was a tie or K<3
      if (k1 < 0 || k2 < 0) tie = true;
    }
    // exact FP32 chains: half 0 rescores k1, half 1 rescores k2
    const long p = p0 + g * 32 + col;
    const int cand = h == 0 ? k1 : k2;
    float s = -INFINITY, xx = 0.0f;
    if (p < M && cand >= 0 && cand < K) {
      const float *x = X + p * D;
      const float *c = C + (long)cand * D;
      s = -0.5f * nc[cand];
      for (int q = 0; q < D; ++q) {
        const float xv = x[q];
        s = __builtin_fmaf(xv, c[q], s);
        xx = __builtin_fmaf(xv, xv, xx);
      }
    }
    const float so = __shfl_xor(s, 32);
    if (h == 0 && p < M) {
      const float s1 = s, s2 = so;  // exact scores of k1, k2
      int lab = k1;
      float sb = s1;
      if (k2 >= 0 && k2 < K && (s2 > s1 || (s2 == s1 && k2 < k1))) { lab = k2; sb = s2; }
      // forward-error bound for every other centre (scores in FP32 units)
      const float xn = sqrtf(xx * (1.0f + 1e-5f)) * (1.0f + 1e-6f) + 1e-30f;
      const float xc = xn * cmax;
      const float Df = (float)D;
      const float g32 = Df * 5.9604645e-8f * 1.0001f;        // gamma_D, u = 2^-24
      const float gmf = (Df + 1.0f) * 2.3841858e-7f * 1.001f; // gamma'_{D+1}, u = 2^-22
      const float rnd = (2.0f * ub + ub * ub) * xc;
      const float etat = eta * sqrtf(Df) * (xn + cmax) + eta * eta * Df;
      const float mfe = gmf * (0.5f * ncmax + (1.0f + ub) * (1.0f + ub) * xc + etat);
      const float fpe = g32 * (0.5f * ncmax + xc);
      const float third = m3 * inv_s2;  // back to FP32 score units
      const float E = (rnd + etat + mfe + fpe) * 1.001f + fabsf(third) * 8e-6f + 1e-30f;
      // no third centre (k <= 2): the two candidates are all there is
      const bool has_third = m3 > -FLT_MAX * 0.5f;
      const bool ok = !tie && (!has_third || sb > third + E);
      if (ok && lab >= 0 && lab < K && sb > -INFINITY) {
        labels[p] = lab;
      } else {
        const int slot = atomicAdd(nflag, 1);
        flags[slot] = (int)p;
      }
    }
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
  hipFree(s->flags);
  hipFree(s->nflag);
  delete s;
  e->mfma = nullptr;
}

void mfma_invalidate(Engine *e) {
  if (st(e)) st(e)->x_ready = false;
}

void mfma_centers_changed(Engine *e, bool new_range) {
  MfmaState *s = st(e);
  if (!s) return;
  s->c_ready = false;
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
  s->Kp = (int)km_cdiv(e->k, MF_TILE) * MF_TILE;
  hipError_t r;
  if ((r = hipMalloc(&s->Xh, sizeof(uint16_t) * (size_t)e->m * s->Dp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->Ch, sizeof(uint16_t) * (size_t)s->Kp * s->Dp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->bias, sizeof(float) * s->Kp)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->bnd, sizeof(float) * 4)) != hipSuccess) return r;
  if ((r = hipMalloc(&s->flags, sizeof(int) * (size_t)e->m)) != hipSuccess) return r;
  s->nflag_slots = 64;
  if ((r = hipMalloc(&s->nflag, sizeof(int) * s->nflag_slots)) != hipSuccess) return r;
  return hipMemsetAsync(s->nflag, 0, sizeof(int) * s->nflag_slots, e->stream);
}

// max |v| over a device array (FP16 scaling only; one host read per run)
__global__ void k_absmax(long n, const float *__restrict__ v, unsigned *__restrict__ out) {
  float m = 0.0f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(v[i]));
  atomicMax(out, __float_as_uint(m));
}

static hipError_t prepare_x(Engine *e) {
  MfmaState *s = st(e);
  if (s->x_ready) return hipSuccess;
  hipError_t r;
  s->scale = 1.0f;
  if (s->fp16) {
    unsigned *dm = nullptr;
    if ((r = hipMalloc(&dm, sizeof(unsigned))) != hipSuccess) return r;
    hipMemsetAsync(dm, 0, sizeof(unsigned), e->stream);
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
  const long tot = (long)e->m * s->Dp;
  hipLaunchKernelGGL(k_convert_rows, dim3(km_cdiv(tot, 256)), dim3(256), 0, e->stream, e->m, e->n,
                     s->Dp, e->dX, s->scale, (int)s->fp16, s->Xh);
  if ((r = hipGetLastError()) != hipSuccess) return r;
  s->x_ready = true;
  return hipSuccess;
}

static hipError_t prepare_centres(Engine *e) {
  MfmaState *s = st(e);
  hipLaunchKernelGGL(k_prepare_centres, dim3(1), dim3(256), 0, e->stream, e->k, e->n, s->Dp, s->Kp,
                     e->dC, e->dnc, s->scale, (int)s->fp16, s->Ch, s->bias, s->bnd);
  s->c_ready = true;
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

template <int KS>
static void launch_mfma_ks(hipStream_t strm, bool fp16, dim3 g, long M, const uint16_t *Xh,
                           const float *X, int D, int K, int Kp, const uint16_t *Ch,
                           const float *bias, const float *C, const float *nc, const float *bnd,
                           float inv_s2, float ub, float eta, int *labels, int *flags, int *nflag) {
  if (fp16)
    hipLaunchKernelGGL((k_assign_mfma<KS, true>), g, dim3(256), 0, strm, M, Xh, X, D, K, Kp, Ch,
                       bias, C, nc, bnd, inv_s2, ub, eta, labels, flags, nflag);
  else
    hipLaunchKernelGGL((k_assign_mfma<KS, false>), g, dim3(256), 0, strm, M, Xh, X, D, K, Kp, Ch,
                       bias, C, nc, bnd, inv_s2, ub, eta, labels, flags, nflag);
}

hipError_t mfma_assign(Engine *e, long r0, long rows) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  if (e->n > 256) return hipErrorNotSupported;
  if (!s->x_ready) {
    if ((r = prepare_x(e)) != hipSuccess) return r;
    s->c_ready = false;
  }
  if (!s->c_ready)
    if ((r = prepare_centres(e)) != hipSuccess) return r;
  if ((r = mfma_ensure_slots(e, e->cur_slot + 1)) != hipSuccess) return r;
  int *nf = s->nflag + e->cur_slot;
  if ((r = hipMemsetAsync(nf, 0, sizeof(int), e->stream)) != hipSuccess) return r;
  if (rows <= 0) return hipSuccess;
  const float ub = s->fp16 ? 4.8828125e-4f : 3.90625e-3f;       // 2^-11, 2^-8
  const float eta = s->fp16 ? 2.9802322e-8f / s->scale : 0.0f;  // half FP16 subnormal step
  const float inv_s2 = 1.0f / (s->scale * s->scale);
  dim3 g(km_cdiv(rows, MF_PTS));
  const uint16_t *Xh = s->Xh + r0 * s->Dp;
  const float *X = e->dX + r0 * e->n;
  int *lab = e->dlabels + r0;
  switch (s->KS) {
    case 1: launch_mfma_ks<1>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
    case 2: launch_mfma_ks<2>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
    case 4: launch_mfma_ks<4>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
    case 6: launch_mfma_ks<6>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
    case 8: launch_mfma_ks<8>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
    default: launch_mfma_ks<16>(e->stream, s->fp16, g, rows, Xh, X, e->n, e->k, s->Kp, s->Ch, s->bias, e->dC, e->dnc, s->bnd, inv_s2, ub, eta, lab, s->flags, nf); break;
  }
  return hipGetLastError();
}

hipError_t mfma_ensure_slots(Engine *e, int n) {
  hipError_t r;
  if ((r = ensure_state(e)) != hipSuccess) return r;
  MfmaState *s = st(e);
  if (n <= s->nflag_slots) return hipSuccess;
  int *nb = nullptr;
  if ((r = hipMalloc(&nb, sizeof(int) * n)) != hipSuccess) return r;
  if ((r = hipMemsetAsync(nb, 0, sizeof(int) * n, e->stream)) != hipSuccess) return r;
  if ((r = hipMemcpyAsync(nb, s->nflag, sizeof(int) * s->nflag_slots, hipMemcpyDeviceToDevice,
                          e->stream)) != hipSuccess)
    return r;
  if ((r = hipStreamSynchronize(e->stream)) != hipSuccess) return r;
  hipFree(s->nflag);
  s->nflag = nb;
  s->nflag_slots = n;
  return hipSuccess;
}

// FP32 full search for the uncertified points (device-side count).
hipError_t mfma_reresolve(Engine *e, long r0, long rows) {
  MfmaState *s = st(e);
  return launch_assign_f32(e->stream, rows, s->nflag + e->cur_slot, e->n, e->dX + r0 * e->n, s->flags,
                           e->k, e->dCt, e->ldc, e->dnc, e->dlabels + r0, nullptr);
}

void mfma_collect(Engine *e, int nslots) {
  MfmaState *s = st(e);
  if (!s || nslots <= 0) return;
  std::vector<int> v(nslots);
  const int n = nslots < s->nflag_slots ? nslots : s->nflag_slots;
  if (hipMemcpyAsync(v.data(), s->nflag, sizeof(int) * n, hipMemcpyDeviceToHost, e->stream) !=
      hipSuccess)
    return;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return;
  for (int i = 0; i < n; ++i) {
    e->stats.reresolved += v[i];
    e->last_reresolved = v[i];
  }
}

}  // namespace km
