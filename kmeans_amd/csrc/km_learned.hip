// Learned candidate proposer (see km_learned.hpp for the method).
//
// Rounding model of the certificates.  The canonical score of centre j is
// s_j = fl(-0.5*nc_j + sum_d x_d c_jd), one fma chain in ascending d, with
// nc_j itself an fma chain of c_jd^2.  With sigma_j = x.c_j - 0.5*||c_j||^2
// (exact), the standard forward-error bound of a length-n fma chain gives
//   |s_j - sigma_j| <= gamma_n (0.5 nc_j + ||x|| ||c_j||) + 0.5 gamma_n ||c_j||^2
//                   <= E(x) = gam * (||x|| cmax + ncmax),
// gam = 1.01 * gamma_{n+2}, gamma_k = k u / (1 - k u), u = 2^-24 (Higham,
// Accuracy and Stability of Numerical Algorithms, 2nd ed., sec. 3.1).  Since
// ||x - c_j||^2 = ||x||^2 - 2 sigma_j, a label j* with score s* is the plain
// FP32 argmax whenever every other centre is shown to have s_j < s*:
//  * separation (Elkan, ICML 2003, lemma 1, restricted to the centres not
//    scored): with a >= ||x - c*|| and h <= 0.5 min ||c* - c_j|| over the
//    members j of the groups not proposed, every such centre has
//    sigma* - sigma_j >= 2 h (h - a); certified when h > a and
//    2 h (h - a) > 2 E (the proposed groups were scored exactly);
//  * group bound: for a group g not proposed, ||x - c_j|| >= t_g =
//    max(0, ||x - G_g|| - r_g) for its members, so s_j <= 0.5 (||x||^2 - t_g^2)
//    + E; certified when s* exceeds that for every such group.
// Squared distances computed as FP32 chains of squared differences are
// within a factor (1 +- gamma_{n+2}) of the exact ones; lower bounds divide by
// (1 + gam), upper bounds multiply by (1 + 2 gam).  All bound arithmetic is
// FP64 with relative slack far above its own rounding.
#include <float.h>
#include <math.h>

#include <algorithm>
#include <chrono>

#include "km_common.hpp"
#include "km_engine.hpp"
#include "km_learned.hpp"
#include "km_mfma.hpp"
#include "km_mlp.hpp"

namespace km {

struct Learned {
  int G = 0, T = 0, H = 0, din = 0, train_iters = 0, proposer = 0;
  long S = 0;            // training rows sampled per plain iteration
  bool ready = false;
  // training samples: row and label of sample i of iteration t at i*train_iters + t
  int *samp_rows = nullptr, *samp_lab = nullptr, *samp_tg = nullptr;
  uint16_t *samp_x = nullptr;
  // groups (membership fixed once built)
  int *gof = nullptr, *members = nullptr, *gstart = nullptr, *gsize = nullptr, *gcol = nullptr;
  int *permP = nullptr;
  int ldP = 0;
  std::vector<int> gsize_h, gcol_h;
  float *CtP = nullptr, *ncP = nullptr, *Gc = nullptr, *gnh = nullptr, *rad2 = nullptr;
  unsigned *sep2 = nullptr;  // k x G: min squared distance from centre i to group g
  int2 *sepTop = nullptr;    // k x (T+1): nearest groups of each centre, ascending
  float *ncmax = nullptr;
  // MLP
  MlpShape sh{};
  size_t Pn = 0;
  float *P = nullptr, *Gr = nullptr, *Am = nullptr, *Av = nullptr, *loss = nullptr;
  uint16_t *Ph = nullptr;
  void *mws = nullptr;
  size_t mws_bytes = 0;
  uint16_t *Xh = nullptr;  // m x din, scaled BF16 inputs
  size_t *dpoff = nullptr;  // device copy of the parameter block offsets
  float xscale = 1.0f;
  long chunk = 0;
  uint16_t *H1 = nullptr, *H2 = nullptr;
  float *Z = nullptr;
  // per-iteration work
  int *cand = nullptr;                      // m x T
  int *gcnt = nullptr, *goff = nullptr, *gcur = nullptr;
  int *ent_pt = nullptr, *ent_id = nullptr, *elab = nullptr;
  float *escore = nullptr;
  int *ul_pt = nullptr, *ul_best = nullptr, *ucnt = nullptr, *fb = nullptr, *fbcnt = nullptr;
  float *ul_score = nullptr;
  unsigned long long *cnt = nullptr;        // separation, group, fallback, proposed, hits
  int *hcnt = nullptr;                      // pinned: gcnt, goff
  double ms_train = 0.0, loss_first = 0.0, loss_last = 0.0;
  long long pairs = 0;                      // exact candidate scores issued
  // grouped MFMA scoring (BF16 engines): group-ordered BF16 centres, biases,
  // per-group wave prefix (pinned host copy + device)
  uint16_t *ChP = nullptr;
  float *biasP = nullptr;
  int *wpre = nullptr, *hwpre = nullptr;
  bool mfma_groups = false;
  // reduced-precision fallback: uncertified rows gathered contiguously
  long fb_cap = 0;
  float *fbX = nullptr;
  float2 *fbN = nullptr;
  uint16_t *fbXh = nullptr;
  int *fbLab = nullptr;
};

static inline Learned *L_of(const Engine *e) { return (Learned *)e->learned; }
static inline unsigned g1(long n) { return (unsigned)((n + 255) / 256); }

static double gamma_n(int n) {
  const double u = ldexp(1.0, -24);
  return 1.01 * ((n + 2) * u) / (1.0 - (n + 2) * u);
}

// ------------------------------------------------------------------ kernels
__global__ void k_fill_i32(long n, int *p, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Workgroup-aggregated counter increment and list append (256 threads, every
// thread of the workgroup calls them): one atomic per workgroup.
__device__ inline void block_count(bool f, unsigned long long *c) {
  __shared__ unsigned bc[4];
  const uint64_t mk = __ballot(f);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) bc[w] = (unsigned)__popcll(mk);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = bc[0] + bc[1] + bc[2] + bc[3];
    if (t) atomicAdd(c, (unsigned long long)t);
  }
  __syncthreads();
}

__device__ inline int block_slot(bool f, int *counter) {
  __shared__ int bs[5];
  const uint64_t mk = __ballot(f);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) bs[w] = __popcll(mk);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = bs[0] + bs[1] + bs[2] + bs[3];
    bs[4] = t ? atomicAdd(counter, t) : 0;
  }
  __syncthreads();
  int base = bs[4];
  for (int ww = 0; ww < w; ++ww) base += bs[ww];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int slot = base + __popcll(mk & below);
  __syncthreads();
  return slot;
}

// out[r][d] = BF16(scale * X[idx ? idx[r] : r][d]) for d < n, 0 for n <= d < din.
__global__ void k_x_bf16(long rows, int n, int din, const float *__restrict__ X,
                         const int *__restrict__ idx, float scale, uint16_t *__restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < rows * din;
       i += (long)gridDim.x * blockDim.x) {
    const long r = i / din;
    const int d = (int)(i - r * din);
    const long src = idx ? (long)idx[r] : r;
    out[i] = mlp_bf16(d < n ? X[src * n + d] * scale : 0.0f);
  }
}

// Mean of q (FP64 squared norms) over sampled rows; one block.
__global__ void k_mean_q(long ns, const int *__restrict__ rows, const double *__restrict__ q,
                         double *__restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = threadIdx.x; i < ns; i += 256) s += q[rows[i]];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0] / (double)(ns > 0 ? ns : 1);
}

__global__ void k_collect(long S, int nit, int it, const int *__restrict__ rows,
                          const int *__restrict__ labels, int *__restrict__ lab) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S) lab[i * nit + it] = labels[rows[i * nit + it]];
}

__global__ void k_targets(long ns, const int *__restrict__ lab, const int *__restrict__ gof,
                          int *__restrict__ tg) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) tg[i] = gof[lab[i]];
}

// Top-T (T <= 8) logits per row; ties to the lower group index.  Sixteen
// lanes per row read the row in 64-byte pieces, keep their own top-T, then T
// rounds of a 16-lane argmax pick the winners.
__global__ __launch_bounds__(256) void k_top_logits(long rows, int G, int T,
                                                    const float *__restrict__ Z,
                                                    int *__restrict__ cand) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long r = gid >> 4;
  const int l = (int)(gid & 15);
  const bool live = r < rows;
  float bv[8];
  int bi[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) { bv[t] = -INFINITY; bi[t] = 0x7fffffff; }
  if (live) {
    const float *z = Z + r * G;
    for (int g = l; g < G; g += 16) {
      float v = z[g];
      int gi = g;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        if (t < T && (v > bv[t] || (v == bv[t] && gi < bi[t]))) {
          const float tv = bv[t];
          const int ti = bi[t];
          bv[t] = v; bi[t] = gi; v = tv; gi = ti;
        }
      }
    }
  }
  for (int t = 0; t < T; ++t) {
    float v = bv[0];
    int i = bi[0];
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
      const float ov = __shfl_xor(v, off);
      const int oi = __shfl_xor(i, off);
      if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
    if (bi[0] == i) {  // this lane's head won: pop it
#pragma unroll
      for (int q = 0; q < 7; ++q) { bv[q] = bv[q + 1]; bi[q] = bi[q + 1]; }
      bv[7] = -INFINITY;
      bi[7] = 0x7fffffff;
    }
    if (live && l == 0) cand[r * T + t] = (v == -INFINITY || i == 0x7fffffff) ? -1 : i;
  }
}

// Nearest-centroid proposer: the T groups with the largest
// x.G_g - 0.5 ||G_g||^2 (empty groups skipped).  One thread per point, the
// point's row in registers (NV float4), centroids staged through LDS.
template <int NV>
__global__ __launch_bounds__(256) void k_group_top(long m, int n, const float *__restrict__ X,
                                                   int G, const float *__restrict__ Gc,
                                                   const float *__restrict__ gnh,
                                                   const int *__restrict__ gsize, int T,
                                                   int *__restrict__ cand) {
  extern __shared__ float4 sg[];  // GT x NV
  const int GT = (int)min((long)G, (long)(65536 / (16 * NV)));
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < m;
  float4 x[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int d = 4 * v + c;
      a[c] = (live && d < n) ? X[p * n + d] : 0.0f;
    }
    x[v] = make_float4(a[0], a[1], a[2], a[3]);
  }
  float bv[8];
  int bi[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) { bv[t] = -INFINITY; bi[t] = -1; }
  for (int g0 = 0; g0 < G; g0 += GT) {
    const int gt = min(GT, G - g0);
    __syncthreads();
    for (int i = threadIdx.x; i < gt * NV; i += blockDim.x) {
      const int gg = i / NV, v = i - gg * NV;
      float a[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int d = 4 * v + c;
        a[c] = d < n ? Gc[(long)(g0 + gg) * n + d] : 0.0f;
      }
      sg[i] = make_float4(a[0], a[1], a[2], a[3]);
    }
    __syncthreads();
    if (!live) continue;
    for (int gg = 0; gg < gt; ++gg) {
      const int g = g0 + gg;
      if (gsize[g] == 0) continue;
      float s = gnh[g];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float4 c = sg[gg * NV + v];
        s = __builtin_fmaf(x[v].x, c.x, s);
        s = __builtin_fmaf(x[v].y, c.y, s);
        s = __builtin_fmaf(x[v].z, c.z, s);
        s = __builtin_fmaf(x[v].w, c.w, s);
      }
      int gi = g;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        if (t < T && s > bv[t]) {
          const float tv = bv[t];
          const int ti = bi[t];
          bv[t] = s; bi[t] = gi; s = tv; gi = ti;
        }
      }
    }
  }
  if (live) {
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (t < T) cand[p * T + t] = bi[t];
  }
}

// Entries e = p*T + t sorted by group: histogram, scan, scatter.
__global__ void k_ent_hist(long ne, const int *__restrict__ cand, int G, int *__restrict__ gcnt) {
  extern __shared__ int h[];
  for (int j = threadIdx.x; j < G; j += blockDim.x) h[j] = 0;
  __syncthreads();
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (long)gridDim.x * blockDim.x) {
    const int g = cand[i];
    if (g >= 0 && g < G) atomicAdd(&h[g], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < G; j += blockDim.x)
    if (h[j]) atomicAdd(&gcnt[j], h[j]);
}

__global__ void k_ent_scan(int G, const int *__restrict__ gcnt, int *__restrict__ goff,
                           int *__restrict__ gcur) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int s = 0;
  for (int g = 0; g < G; ++g) {
    goff[g] = s;
    gcur[g] = s;
    s += gcnt[g];
  }
}

#define ENT_CHUNK 8192
__global__ void k_ent_scatter(long ne, int T, const int *__restrict__ cand, int G,
                              int *__restrict__ gcur, int *__restrict__ ent_pt,
                              int *__restrict__ ent_id) {
  extern __shared__ int sh[];  // h[G], base[G]
  int *h = sh, *base = sh + G;
  const long a = (long)blockIdx.x * ENT_CHUNK, e = min(a + ENT_CHUNK, ne);
  for (int j = threadIdx.x; j < G; j += blockDim.x) h[j] = 0;
  __syncthreads();
  for (long i = a + threadIdx.x; i < e; i += blockDim.x) {
    const int g = cand[i];
    if (g >= 0 && g < G) atomicAdd(&h[g], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < G; j += blockDim.x) {
    base[j] = h[j] ? atomicAdd(&gcur[j], h[j]) : 0;
    h[j] = 0;
  }
  __syncthreads();
  for (long i = a + threadIdx.x; i < e; i += blockDim.x) {
    const int g = cand[i];
    if (g < 0 || g >= G) continue;
    const int pos = base[g] + atomicAdd(&h[g], 1);
    ent_pt[pos] = (int)(i / T);
    ent_id[pos] = (int)i;
  }
}

// Centre columns in group order: CtP[d][c] = C[permP[c]][d] (0 for padding).
__global__ void k_gather_perm(int ldP, int n, const int *__restrict__ permP,
                              const float *__restrict__ C, const float *__restrict__ nc,
                              float *__restrict__ CtP, float *__restrict__ ncP) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)ldP * n) return;
  const int d = (int)(i / ldP), c = (int)(i - (long)d * ldP);
  const int j = permP[c];
  CtP[i] = j >= 0 ? C[(long)j * n + d] : 0.0f;
  if (d == 0) ncP[c] = j >= 0 ? nc[j] : 0.0f;
}

// Per group: centroid (FP64 mean of the members, stored FP32), -0.5||G||^2,
// and rad2 = max over members of the FP32 squared distance to the centroid.
__global__ __launch_bounds__(256) void k_group_geom(int n, const int *__restrict__ gstart,
                                                    const int *__restrict__ members,
                                                    const float *__restrict__ C,
                                                    float *__restrict__ Gc,
                                                    float *__restrict__ gnh,
                                                    float *__restrict__ rad2) {
  extern __shared__ float gs[];  // n
  __shared__ float red[256];
  const int g = blockIdx.x, a = gstart[g], b = gstart[g + 1], cnt = b - a;
  for (int d = threadIdx.x; d < n; d += blockDim.x) {
    double s = 0.0;
    for (int i = a; i < b; ++i) s += (double)C[(long)members[i] * n + d];
    const float v = cnt ? (float)(s / (double)cnt) : 0.0f;
    gs[d] = v;
    Gc[(long)g * n + d] = v;
  }
  __syncthreads();
  float mx = 0.0f;
  for (int i = a + threadIdx.x; i < b; i += blockDim.x) {
    const float *c = C + (long)members[i] * n;
    float dd = 0.0f;
    for (int d = 0; d < n; ++d) {
      const float df = c[d] - gs[d];
      dd = __builtin_fmaf(df, df, dd);
    }
    mx = fmaxf(mx, dd);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rad2[g] = red[0];
    float s = 0.0f;
    for (int d = 0; d < n; ++d) s = __builtin_fmaf(gs[d], gs[d], s);
    gnh[g] = -0.5f * s;
  }
}

// D[i][g] = min over members j != i of group g of the FP32 squared distance
// ||c_i - c_j||^2 (chain of squared differences, ascending d), as float bits;
// D must hold +inf bits on entry.  Rows i in centre order, columns in group
// order (CtP), so a 64-column tile spans few groups: minima are first taken
// per (row, group) in LDS, then one global atomicMin per (row, group).
#define SEP_T 64
#define SEP_DS 32
__global__ __launch_bounds__(256) void k_center_sep(int k, int n, int G, const float *__restrict__ C,
                                                    const float *__restrict__ CtP, int ldP,
                                                    const int *__restrict__ permP,
                                                    const int *__restrict__ gof,
                                                    unsigned *__restrict__ D) {
  __shared__ float A[SEP_DS][SEP_T + 1], B[SEP_DS][SEP_T + 1];
  __shared__ unsigned mins[SEP_T][SEP_T];
  __shared__ int colj[SEP_T], colg[SEP_T], colslot[SEP_T], slotg[SEP_T], nslot;
  const int i0 = blockIdx.y * SEP_T, c0 = blockIdx.x * SEP_T;
  const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
  if (tid < SEP_T) {
    const int c = c0 + tid;
    const int j = c < ldP ? permP[c] : -1;
    colj[tid] = j;
    colg[tid] = j >= 0 ? gof[j] : -1;
  }
  for (int e = tid; e < SEP_T * SEP_T; e += 256) (&mins[0][0])[e] = 0x7f800000u;
  __syncthreads();
  if (tid == 0) {  // columns are in group order: one slot per distinct group
    int ns = 0, last = -1;
    for (int q = 0; q < SEP_T; ++q) {
      const int g = colg[q];
      if (g >= 0 && g != last) { slotg[ns++] = g; last = g; }
      colslot[q] = ns - 1;
    }
    nslot = ns;
  }
  float acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
  for (int d0 = 0; d0 < n; d0 += SEP_DS) {
    __syncthreads();
    for (int e = tid; e < SEP_T * SEP_DS; e += 256) {
      const int r = e / SEP_DS, dd = e - r * SEP_DS;
      const int d = d0 + dd;
      A[dd][r] = (i0 + r < k && d < n) ? C[(long)(i0 + r) * n + d] : 0.0f;
    }
    for (int e = tid; e < SEP_T * SEP_DS; e += 256) {
      const int dd = e / SEP_T, r = e - dd * SEP_T;
      const int d = d0 + dd;
      B[dd][r] = (c0 + r < ldP && d < n) ? CtP[(long)d * ldP + c0 + r] : 0.0f;
    }
    __syncthreads();
    const int dn = min(SEP_DS, n - d0);
    for (int dd = 0; dd < dn; ++dd) {
      float av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = A[dd][ti * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = B[dd][tj + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float df = av[a] - bv[b];
          acc[a][b] = __builtin_fmaf(df, df, acc[a][b]);
        }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ti * 4 + a;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int q = tj + 16 * b;
      const int j = colj[q];
      if (i < k && j >= 0 && j != i)
        atomicMin(&mins[ti * 4 + a][colslot[q]], __float_as_uint(acc[a][b]));
    }
  }
  __syncthreads();
  const int ns = nslot;
  for (int e = tid; e < SEP_T * ns; e += 256) {
    const int r = e / ns, sl = e - r * ns;
    const int i = i0 + r;
    const unsigned v = mins[r][sl];
    if (i < k && v != 0x7f800000u) atomicMin(&D[(long)i * G + slotg[sl]], v);
  }
}

// Per centre: the NT smallest entries of its row of D (squared separation to
// each group, float bits) with their groups, ascending, as (bits, group);
// missing entries are (+inf, -1).  The nearest centre outside any set of at
// most NT - 1 named groups is then the first listed group not named.
__global__ void k_sep_top(int k, int G, int NT, const unsigned *__restrict__ D,
                          int2 *__restrict__ top) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  unsigned v[9];
  int gi[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { v[i] = 0x7f800000u; gi[i] = -1; }
  const unsigned *row = D + (long)c * G;
  for (int g = 0; g < G; ++g) {
    unsigned x = row[g];
    int xg = g;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      if (i < NT && x < v[i]) {
        const unsigned tv = v[i];
        const int tg = gi[i];
        v[i] = x; gi[i] = xg; x = tv; xg = tg;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i)
    if (i < NT) top[(long)c * NT + i] = make_int2((int)v[i], gi[i]);
}

__global__ void k_ncmax(int k, const float *__restrict__ nc, float *__restrict__ out) {
  __shared__ float red[256];
  float mx = 0.0f;
  for (int j = threadIdx.x; j < k; j += 256) mx = fmaxf(mx, nc[j]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

struct BoundParams {
  double gam;  // 1.01 gamma_{n+2}
};

__device__ inline double score_err(double qu, double ncm, double gam) {
  const double cmax = sqrt(ncm * (1.0 + gam)) * (1.0 + 1e-15);
  return gam * (sqrt(qu) * cmax + ncm) * (1.0 + 1e-12) + 1e-300;
}

// Best candidate per point (largest canonical score, lowest original index on
// ties) and the separation certificate; the rest go to the uncertified list.
__global__ __launch_bounds__(256) void k_combine(
    long m, int T, const int *__restrict__ cand, const int *__restrict__ gsize,
    const int *__restrict__ gcol, const int *__restrict__ permP, const int *__restrict__ elab,
    const float *__restrict__ escore, const double *__restrict__ q,
    const int2 *__restrict__ sepTop, const float *__restrict__ ncmax, BoundParams bp,
    int *__restrict__ labels, int *__restrict__ ul_pt, int *__restrict__ ul_best,
    float *__restrict__ ul_score, int *__restrict__ ucnt, unsigned long long *__restrict__ cnt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < m;
  int bj = -1;
  float bs = -INFINITY;
  bool cert = false;
  if (live) {
    bool unsure = false;  // a named group whose best member was not certified
    for (int t = 0; t < T; ++t) {
      const long e = p * T + t;
      const int g = cand[e];
      if (g < 0 || gsize[g] == 0) continue;
      const int le = elab[e];
      if (le < 0) { unsure = true; continue; }
      const int j = permP[gcol[g] + le];
      const float s = escore[e];
      if (s > bs || (s == bs && j < bj)) { bs = s; bj = j; }
    }
    if (unsure) bj = -1;
    if (bj >= 0 && bs > -INFINITY) {
      const double qu = q[p] * (1.0 + 1e-12);
      const double E = score_err(qu, (double)ncmax[0], bp.gam);
      const double a2 = qu - 2.0 * (double)bs + 2.0 * E;
      const double a = sqrt(fmax(a2, 0.0)) * (1.0 + 1e-14);
      // nearest centre outside the named groups (their members were scored):
      // the first of bj's T+1 nearest groups that is not named
      const int2 *tl = sepTop + (long)bj * (T + 1);
      float s2f = INFINITY;
      for (int i = 0; i <= T; ++i) {
        const int2 en = tl[i];
        if (en.y < 0) break;
        bool named = false;
        for (int t = 0; t < T; ++t) named |= cand[p * T + t] == en.y;
        if (!named) { s2f = __uint_as_float((unsigned)en.x); break; }
      }
      const double s2 = (double)s2f;
      const double h = s2f == INFINITY ? INFINITY : 0.5 * sqrt(s2 / (1.0 + bp.gam)) * (1.0 - 1e-14);
      cert = (h > a) && (2.0 * h * (h - a) > 2.0 * E * (1.0 + 1e-12));
    }
  }
  if (cert) labels[p] = bj;
  const bool unc = live && !cert;
  const int slot = block_slot(unc, ucnt);
  if (unc) {
    ul_pt[slot] = (int)p;
    ul_best[slot] = bj;
    ul_score[slot] = bs;
  }
  block_count(cert, &cnt[0]);
  block_count(live, &cnt[3]);
}

// Group bound for the uncertified points; the rest go to the fallback list.
__global__ __launch_bounds__(256) void k_group_bound(
    const int *__restrict__ ucnt, int n, const float *__restrict__ X, int T,
    const int *__restrict__ cand, int G, const float *__restrict__ Gc,
    const float *__restrict__ rad2, const int *__restrict__ gsize, const double *__restrict__ q,
    const float *__restrict__ ncmax, BoundParams bp, const int *__restrict__ ul_pt,
    const int *__restrict__ ul_best, const float *__restrict__ ul_score, int *__restrict__ labels,
    int *__restrict__ fb, int *__restrict__ fbcnt, unsigned long long *__restrict__ cnt) {
  const long nu = *ucnt;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long b0 = (long)blockIdx.x * blockDim.x; b0 < nu; b0 += stride) {
    const long i = b0 + threadIdx.x;
    const bool live = i < nu;
    bool cert = false;
    int p = 0;
    if (live) {
      p = ul_pt[i];
      const int bj = ul_best[i];
      const double bs = (double)ul_score[i];
      if (bj >= 0) {
        const double qu = q[p] * (1.0 + 1e-12);
        const double E = score_err(qu, (double)ncmax[0], bp.gam);
        const float *x = X + (long)p * n;
        cert = true;
        for (int g = 0; g < G && cert; ++g) {
          if (gsize[g] == 0) continue;
          bool named = false;
          for (int t = 0; t < T; ++t) named |= cand[(long)p * T + t] == g;
          if (named) continue;
          const float *c = Gc + (long)g * n;
          float dd = 0.0f;
          for (int d = 0; d < n; ++d) {
            const float df = x[d] - c[d];
            dd = __builtin_fmaf(df, df, dd);
          }
          const double dlow = sqrt((double)dd / (1.0 + bp.gam)) * (1.0 - 1e-14);
          const double r = sqrt((double)rad2[g] * (1.0 + 2.0 * bp.gam)) * (1.0 + 1e-14);
          const double t = fmax(dlow - r, 0.0);
          const double ub = 0.5 * (qu - t * t);
          cert = bs > ub + E * (1.0 + 1e-12);
        }
      }
    }
    if (cert) labels[p] = ul_best[i];
    const bool f = live && !cert;
    const int slot = block_slot(f, fbcnt);
    if (f) fb[slot] = p;
    block_count(cert, &cnt[1]);
    block_count(f, &cnt[2]);
  }
}

// Proposal hits: points whose final label lies in one of their named groups.
__global__ void k_hits(long m, int T, const int *__restrict__ cand, const int *__restrict__ gof,
                       const int *__restrict__ labels, unsigned long long *__restrict__ cnt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (p < m) {
    const int g = gof[labels[p]];
    for (int t = 0; t < T; ++t) hit |= cand[p * T + t] == g;
  }
  block_count(hit, &cnt[4]);
}

// Group-ordered BF16 centre rows and biases for the grouped MFMA scoring:
// row c of ChP is row permP[c] of Ch (zeros for padding).
__global__ void k_gather_ch(int ldP, int Dp, const int *__restrict__ permP,
                            const uint16_t *__restrict__ Ch, const float *__restrict__ bias,
                            uint16_t *__restrict__ ChP, float *__restrict__ biasP) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // one 16-byte chunk
  const int cpr = Dp / 8;
  if (i >= (long)ldP * cpr) return;
  const int c = (int)(i / cpr), q = (int)(i - (long)c * cpr);
  const int j = permP[c];
  typedef __attribute__((ext_vector_type(4))) unsigned int u4;
  u4 v = {0u, 0u, 0u, 0u};
  if (j >= 0) v = *reinterpret_cast<const u4 *>(Ch + (long)j * Dp + q * 8);
  *reinterpret_cast<u4 *>(ChP + (long)c * Dp + q * 8) = v;
  if (q == 0) biasP[c] = j >= 0 ? bias[j] : -FLT_MAX;
}

typedef __attribute__((ext_vector_type(8))) __bf16 lg_bf16x8;
typedef __attribute__((ext_vector_type(16))) float lg_f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int lg_u32x4;
__device__ inline int lg_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// Grouped candidate scoring on BF16 MFMA.  One wave scores 32 entries of
// one group (entries sorted by group; wave -> group by the prefix wpre of
// ceil(count/32)) against every member of the group: BF16 scores with FP32
// accumulation, a top-3 per entry, the exact FP32 chains of the first two
// (half 0 and half 1 of the wave), and the within-group certificate: the
// better exact score (lowest index on ties) exceeds the third approximate
// score + E, E being the BF16 forward-error bound of the plain path's
// certificate.  Writes the exact score and the member's local index, or -1
// when the group's best member is not certified (its point is then searched
// in full).
template <int KS>
__global__ __launch_bounds__(256) void k_group_mfma(
    int G, const int *__restrict__ wpre, const int *__restrict__ goff,
    const int *__restrict__ gcnt, const int *__restrict__ gsize, const int *__restrict__ gcol,
    const int *__restrict__ permP, const int *__restrict__ ent_pt,
    const int *__restrict__ ent_id, const uint16_t *__restrict__ Xh,
    const float *__restrict__ X, int D, const uint16_t *__restrict__ ChP,
    const float *__restrict__ biasP, const float *__restrict__ C, const float *__restrict__ nc,
    const float *__restrict__ bnd, const float2 *__restrict__ xnb, int *__restrict__ elab,
    float *__restrict__ escore, float big) {
  constexpr int Dp = KS * 16;
  constexpr int CPR = Dp / 8;                    // 16-byte chunks per centre row
  constexpr int ROTM = CPR >= 16 ? 15 : 7;       // LDS swizzle rotation mask
  __shared__ lg_u32x4 ts[2][32 * CPR];           // two centre tiles (double buffer)
  __shared__ float tb[2][32];
  __shared__ int wg[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
  const int wv = blockIdx.x * 4 + w;
  const int nw = wpre[G];
  // group of this wave (-1 past the end)
  int g = -1;
  if (wv < nw) {
    int lo = 0, hi = G - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (wpre[mid] <= wv) lo = mid; else hi = mid - 1;
    }
    g = lo;
  }
  if (lane == 0) wg[w] = g;
  __syncthreads();
  // the workgroup shares LDS tiles when all its live waves score one group
  const int g0 = wg[0];
  const bool shared = g0 >= 0 && (wg[1] == g0 || wg[1] < 0) && (wg[2] == g0 || wg[2] < 0) &&
                      (wg[3] == g0 || wg[3] < 0);
  if (g < 0 && !shared) return;  // idle wave of a mixed workgroup (no barriers follow)
  const int gg = shared ? g0 : g;
  const int eb = goff[gg], ee = eb + gcnt[gg];
  const int ei = g >= 0 ? eb + (wv - wpre[gg]) * 32 + col : ee;
  const bool ok = ei < ee;
  const long p = ok ? (long)ent_pt[ei] : 0;
  lg_u32x4 bfr[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
    bfr[ks] = ok ? *reinterpret_cast<const lg_u32x4 *>(Xh + p * Dp + ks * 16 + 8 * h)
                 : lg_u32x4{0u, 0u, 0u, 0u};
  const int sz = gsize[gg], c0 = gcol[gg];
  const int nt = (sz + 31) / 32;
  // top 3 of packed scores (register id in the low 4 bits) and tiles of the first two
  float b1 = -FLT_MAX, b2 = -FLT_MAX, b3 = -FLT_MAX;
  int t1 = 0, t2 = 0;
  auto epilogue = [&](int t, const lg_f32x16 &acc) {
    const float o1 = b1, o2 = b2;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float pv = __uint_as_float((__float_as_uint(fmaxf(acc[r], -FLT_MAX)) & 0xFFFFFFF0u) | (unsigned)r);
      const float n3 = __builtin_amdgcn_fmed3f(b2, b3, pv);
      const float n2 = __builtin_amdgcn_fmed3f(b1, b2, pv);
      b1 = __builtin_amdgcn_fmed3f(b1, pv, big);
      b2 = n2;
      b3 = n3;
    }
    const unsigned u1 = __float_as_uint(b1), u2 = __float_as_uint(b2);
    const int ot1 = t1, ot2 = t2;
    t2 = (u2 == __float_as_uint(o1)) ? ot1 : (u2 == __float_as_uint(o2)) ? ot2 : t;
    t1 = (u1 != __float_as_uint(o1)) ? t : ot1;
  };
  auto tile_bias = [&](int t, const float *bsrc, lg_f32x16 &acc) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int lr = t * 32 + lg_row(r, h);
      acc[r] = lr < sz ? bsrc[lg_row(r, h)] : -INFINITY;
    }
  };
  if (shared) {
    // stage tile t into buffer t & 1 (rows past the group's end read the
    // padding/next group's rows, masked by the bias)
    auto stage = [&](int t) {
      const int b = t & 1;
      for (int c = threadIdx.x; c < 32 * CPR; c += 256) {
        const int r = c / CPR, q = c - r * CPR;
        ts[b][r * CPR + (q + (r & ROTM)) % CPR] =
            *reinterpret_cast<const lg_u32x4 *>(ChP + (long)(c0 + t * 32 + r) * Dp + q * 8);
      }
      if (threadIdx.x < 32) tb[b][threadIdx.x] = biasP[c0 + t * 32 + threadIdx.x];
    };
    stage(0);
    for (int t = 0; t < nt; ++t) {
      __syncthreads();
      if (t + 1 < nt) stage(t + 1);
      const int b = t & 1;
      lg_f32x16 acc;
      tile_bias(t, tb[b], acc);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int q = ks * 2 + h;
        const lg_u32x4 a = ts[b][col * CPR + (q + (col & ROTM)) % CPR];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a),
                                                       __builtin_bit_cast(lg_bf16x8, bfr[ks]), acc, 0, 0, 0);
      }
      epilogue(t, acc);
    }
  } else {
    for (int t = 0; t < nt; ++t) {
      lg_f32x16 acc;
      tile_bias(t, biasP + c0 + t * 32, acc);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const lg_u32x4 a =
            *reinterpret_cast<const lg_u32x4 *>(ChP + (long)(c0 + t * 32 + col) * Dp + ks * 16 + 8 * h);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a),
                                                       __builtin_bit_cast(lg_bf16x8, bfr[ks]), acc, 0, 0, 0);
      }
      epilogue(t, acc);
    }
  }
  if (g < 0) return;  // idle wave of a shared workgroup (after the barriers)
  // local indices of the first two; bitwise-equal packed values are ties
  const bool real1 = b1 > -FLT_MAX, real2 = b2 > -FLT_MAX;
  int i1 = real1 ? t1 * 32 + lg_row((int)(__float_as_uint(b1) & 15u), h) : -1;
  int i2 = real2 ? t2 * 32 + lg_row((int)(__float_as_uint(b2) & 15u), h) : -1;
  bool tie = (__float_as_uint(b1) == __float_as_uint(b2)) || (__float_as_uint(b2) == __float_as_uint(b3));
  tie = tie || (__shfl_xor((int)tie, 32) != 0);
  {
    const float w1 = __shfl_xor(b1, 32), w2 = __shfl_xor(b2, 32), w3 = __shfl_xor(b3, 32);
    const int j1 = __shfl_xor(i1, 32), j2 = __shfl_xor(i2, 32);
    const bool A1 = b1 >= w1;
    const float m1 = A1 ? b1 : w1;
    const int k1 = A1 ? i1 : j1;
    const float x2 = A1 ? b2 : b1, y2 = A1 ? w1 : w2;
    const int xi2 = A1 ? i2 : i1, yi2 = A1 ? j1 : j2;
    const bool A2 = x2 >= y2;
    const float m2 = A2 ? x2 : y2;
    const int k2 = A2 ? xi2 : yi2;
    const float xa = A1 ? (A2 ? b3 : b2) : (A2 ? b2 : b1);
    const float ya = A1 ? (A2 ? w1 : w2) : (A2 ? w2 : w3);
    const float m3 = fmaxf(xa, ya);
    const bool mt = (__float_as_uint(m1) == __float_as_uint(m2)) || (__float_as_uint(m2) == __float_as_uint(m3));
    b1 = __shfl(m1, col);
    b2 = __shfl(m2, col);
    b3 = __shfl(m3, col);
    i1 = __shfl(k1, col);
    i2 = __shfl(k2, col);
    tie = tie || (__shfl((int)mt, col) != 0);
  }
  // exact chains: half 0 for the first candidate, half 1 for the second
  const int lc = h ? i2 : i1;
  const int jc = (ok && lc >= 0 && lc < sz) ? permP[c0 + lc] : -1;
  float sx = -INFINITY;
  if (jc >= 0) {
    const float *xr = X + p * D;
    const float *cr = C + (long)jc * D;
    float sc = -0.5f * nc[jc];
    if ((D & 3) == 0) {
      const float4 *x4 = reinterpret_cast<const float4 *>(xr);
      const float4 *c4 = reinterpret_cast<const float4 *>(cr);
#pragma unroll 4
      for (int q = 0; q < D / 4; ++q) {
        const float4 xv = x4[q], cv = c4[q];
        sc = __builtin_fmaf(xv.x, cv.x, sc);
        sc = __builtin_fmaf(xv.y, cv.y, sc);
        sc = __builtin_fmaf(xv.z, cv.z, sc);
        sc = __builtin_fmaf(xv.w, cv.w, sc);
      }
    } else {
      for (int q = 0; q < D; ++q) sc = __builtin_fmaf(xr[q], cr[q], sc);
    }
    sx = sc;
  }
  const float sy = __shfl_xor(sx, 32);
  const int jy = __shfl_xor(jc, 32);
  if (ok && h == 0) {
    int lab = i1, jb = jc;
    float sb = sx;
    if (jy >= 0 && (sy > sb || (sy == sb && jy < jb))) { sb = sy; lab = i2; jb = jy; }
    // the plain path's bound: |x| and |x - xh| per point, max |c_j|,
    // |c_j|^2 and |c_j - ch_j| over the centres (BF16, no scale)
    const float cmax = bnd[0], ncmax = bnd[1], dcmax = bnd[2];
    const float2 nb = xnb[p];
    const float xn = nb.x, ex = nb.y, Df = (float)D;
    const float g32 = Df * 5.9604645e-8f * 1.0001f;                // gamma_D, u = 2^-24
    const float gmf = (Df + 1.0f) * 2.3841858e-7f * 1.001f;        // gamma'_{D+1}, u = 2^-22
    const float rnd = ex * cmax + (xn + ex) * dcmax;
    const float mfe = gmf * (0.5f * ncmax + (xn + ex) * (cmax + dcmax));
    const float fpe = g32 * (0.5f * ncmax + xn * cmax);
    const float E = (rnd + mfe + fpe) * 1.001f + 1e-30f;
    const bool has3 = b3 > -FLT_MAX;
    // packed low bits (< 16 ulp) and the FP32 comparison
    const float slack = (fabsf(b1) + fabsf(b2) + (has3 ? fabsf(b3) : 0.0f)) * 1.6e-5f;
    const float thr = has3 ? b3 + E + slack : -INFINITY;
    const bool cert = !tie && jb >= 0 && sb > -INFINITY && sb > thr;
    const int slot = ent_id[ei];
    elab[slot] = cert ? lab : -1;
    escore[slot] = sb;
  }
}

// Fused MLP inference and top-T proposal.  One workgroup (4 waves) takes 64
// rows of the scaled BF16 inputs: layer 1 (B fragments straight from the
// inputs), layer 2 and layer 3 on v_mfma_f32_32x32x16_bf16 with the hidden
// activations kept in LDS as BF16 (row-major, padded rows), weights read from
// the (L2-resident) BF16 parameter copy, bias as the accumulator input, ReLU
// on the way into LDS.  Layer 3 is done by waves 0 and 1 (32 rows each, all G
// logits): every lane keeps the top T of its half of the logits, the halves
// are merged, and only the T group ids are written.  Requires din, H and G
// multiples of 32 (din of 16), G <= 256, T <= 8.
#define MLPF_ROWS 64
__device__ inline void mlpf_store4(uint16_t *row, int n0, const float *v) {
  // four consecutive neurons n0..n0+3 as BF16, ReLU applied
  const uint32_t a = bf16x2_rne(fmaxf(v[0], 0.0f), fmaxf(v[1], 0.0f));
  const uint32_t b = bf16x2_rne(fmaxf(v[2], 0.0f), fmaxf(v[3], 0.0f));
  *reinterpret_cast<uint2 *>(row + n0) = make_uint2(a, b);
}

template <int KH>
__global__ __launch_bounds__(256) void k_mlp_propose(long rows, int din, int G, int T,
                                                     const uint16_t *__restrict__ Xh,
                                                     const uint16_t *__restrict__ Ph,
                                                     const float *__restrict__ P,
                                                     const size_t *__restrict__ poff,
                                                     int *__restrict__ cand) {
  constexpr int H = KH * 16;       // hidden width
  constexpr int NTW = H / 32 / 4;  // neuron tiles per wave (layers 1, 2)
  extern __shared__ uint16_t hs[];  // H1 then H2: MLPF_ROWS x (H + 8) each; logits reuse H1
  constexpr int ldh = H + 8;
  uint16_t *H1 = hs, *H2 = hs + MLPF_ROWS * ldh;
  float *Zs = reinterpret_cast<float *>(hs);  // MLPF_ROWS x (G + 4), after layer 2
  const int ldz = G + 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
  const long r0 = (long)blockIdx.x * MLPF_ROWS;
  const uint16_t *W1 = Ph + poff[0], *W2 = Ph + poff[2], *W3 = Ph + poff[4];
  const float *b1 = P + poff[1], *b2 = P + poff[3], *b3 = P + poff[5];
  const long pr0 = r0 + col, pr1 = r0 + 32 + col;  // this lane's rows in the two row tiles
  const bool ok0 = pr0 < rows, ok1 = pr1 < rows;
  // one layer: out[rows][nt*32..] = relu(in[rows][:] W^T + b) for this wave's
  // neuron tiles and both row tiles; B fragments from LDS (or the inputs)
  auto store_tile = [&](uint16_t *dst_rows, int nt, int rt, const lg_f32x16 &acc) {
    uint16_t *dst = dst_rows + (rt * 32 + col) * ldh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float v[4] = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
      mlpf_store4(dst, nt * 32 + 8 * q + 4 * h, v);
    }
  };
  // ---- layer 1 (K = din, runtime)
#pragma unroll
  for (int u = 0; u < NTW; ++u) {
    const int nt = w * NTW + u;
    lg_f32x16 a0, a1;
#pragma unroll
    for (int r = 0; r < 16; ++r) a0[r] = a1[r] = b1[nt * 32 + lg_row(r, h)];
    for (int ks = 0; ks < din / 16; ++ks) {
      const lg_u32x4 a = *reinterpret_cast<const lg_u32x4 *>(W1 + (long)(nt * 32 + col) * din + ks * 16 + 8 * h);
      const lg_u32x4 x0 = ok0 ? *reinterpret_cast<const lg_u32x4 *>(Xh + pr0 * din + ks * 16 + 8 * h)
                              : lg_u32x4{0u, 0u, 0u, 0u};
      const lg_u32x4 x1 = ok1 ? *reinterpret_cast<const lg_u32x4 *>(Xh + pr1 * din + ks * 16 + 8 * h)
                              : lg_u32x4{0u, 0u, 0u, 0u};
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x0), a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x1), a1, 0, 0, 0);
    }
    store_tile(H1, nt, 0, a0);
    store_tile(H1, nt, 1, a1);
  }
  __syncthreads();
  // ---- layer 2 (K = H, unrolled: the weight loads overlap)
#pragma unroll
  for (int u = 0; u < NTW; ++u) {
    const int nt = w * NTW + u;
    lg_f32x16 a0, a1;
#pragma unroll
    for (int r = 0; r < 16; ++r) a0[r] = a1[r] = b2[nt * 32 + lg_row(r, h)];
    const uint16_t *s0 = H1 + col * ldh, *s1 = H1 + (32 + col) * ldh;
#pragma unroll
    for (int ks = 0; ks < KH; ++ks) {
      const lg_u32x4 a = *reinterpret_cast<const lg_u32x4 *>(W2 + (long)(nt * 32 + col) * H + ks * 16 + 8 * h);
      const lg_u32x4 x0 = *reinterpret_cast<const lg_u32x4 *>(s0 + ks * 16 + 8 * h);
      const lg_u32x4 x1 = *reinterpret_cast<const lg_u32x4 *>(s1 + ks * 16 + 8 * h);
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x0), a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x1), a1, 0, 0, 0);
    }
    store_tile(H2, nt, 0, a0);
    store_tile(H2, nt, 1, a1);
  }
  __syncthreads();
  // ---- layer 3: logits of this wave's group tiles into LDS (over H1)
  for (int nt = w; nt < G / 32; nt += 4) {
    lg_f32x16 a0, a1;
#pragma unroll
    for (int r = 0; r < 16; ++r) a0[r] = a1[r] = b3[nt * 32 + lg_row(r, h)];
    const uint16_t *s0 = H2 + col * ldh, *s1 = H2 + (32 + col) * ldh;
#pragma unroll
    for (int ks = 0; ks < KH; ++ks) {
      const lg_u32x4 a = *reinterpret_cast<const lg_u32x4 *>(W3 + (long)(nt * 32 + col) * H + ks * 16 + 8 * h);
      const lg_u32x4 x0 = *reinterpret_cast<const lg_u32x4 *>(s0 + ks * 16 + 8 * h);
      const lg_u32x4 x1 = *reinterpret_cast<const lg_u32x4 *>(s1 + ks * 16 + 8 * h);
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x0), a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, a), __builtin_bit_cast(lg_bf16x8, x1), a1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      Zs[col * ldz + nt * 32 + lg_row(r, h)] = a0[r];
      Zs[(32 + col) * ldz + nt * 32 + lg_row(r, h)] = a1[r];
    }
  }
  __syncthreads();
  // ---- top T per row (ties to the lower group index)
  if (threadIdx.x < MLPF_ROWS) {
    const long pr = r0 + threadIdx.x;
    if (pr < rows) {
      float bv[8];
      int bi[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) { bv[t] = -INFINITY; bi[t] = 0x7fffffff; }
      const float *z = Zs + threadIdx.x * ldz;
      for (int g = 0; g < G; ++g) {
        float v = z[g];
        int gi = g;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          if (t < T && v > bv[t]) {
            const float tv = bv[t];
            const int ti = bi[t];
            bv[t] = v; bi[t] = gi; v = tv; gi = ti;
          }
        }
      }
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (t < T) cand[pr * T + t] = (bv[t] == -INFINITY) ? -1 : bi[t];
    }
  }
}

// Weight-stationary fused MLP inference and top-T proposal (persistent: one
// workgroup of H/32 waves per CU walks 64-row blocks).  Wave w owns neuron
// tile w of layers 1 and 2: its W1 and W2 fragments and biases stay in
// registers for the whole kernel; W3 is copied once into LDS.  Per block:
// layer 1 from the rows' BF16 fragments (read one block ahead into
// registers), layer 2 from H1 in LDS, layer 3 by (group tile, row tile)
// pairs from H2 and W3 in LDS into FP32 logits, then the top T of each row
// by H/32 lanes: each logit carries its group in the low 7 bits (G - 1 - g
// for non-negative values, g for negative ones, so equal logits order by
// the lower group index), the lanes keep a top-TT by a v_med3/v_max network
// over strided groups and merge by shuffles.  The products, accumulation
// order and roundings are those of k_mlp_propose; the packed bits move a
// logit by less than 2^-16 of its value.  Requires din = 16*KI, G % 32 ==
// 0, G <= 128, T <= TT <= 8.
#define MLPW_ROWS 64
__host__ __device__ constexpr int mlpw_ldh(int H) { return H + 8; }
__host__ __device__ inline size_t mlpw_zs_off(int H, int G) {
  // logits over H1 when they fit there, else after H2 (in uint16 units)
  return (size_t)MLPW_ROWS * (G + 4) * 4 <= (size_t)MLPW_ROWS * mlpw_ldh(H) * 2
             ? (size_t)G * mlpw_ldh(H)
             : (size_t)G * mlpw_ldh(H) + 2 * (size_t)MLPW_ROWS * mlpw_ldh(H);
}
__host__ __device__ inline size_t mlpw_lds_bytes(int H, int G) {
  const size_t zs_end = mlpw_zs_off(H, G) * 2 + (size_t)MLPW_ROWS * (G + 4) * 4;
  const size_t h_end = ((size_t)G * mlpw_ldh(H) + 2 * (size_t)MLPW_ROWS * mlpw_ldh(H)) * 2;
  return zs_end > h_end ? zs_end : h_end;
}
template <int KH, int KI, int TT>
__global__ __launch_bounds__(KH / 2 * 64) void k_mlp_propose_ws(long rows, int G, int T,
                                                              const uint16_t *__restrict__ Xh,
                                                              const uint16_t *__restrict__ Ph,
                                                              const float *__restrict__ P,
                                                              const size_t *__restrict__ poff,
                                                              int *__restrict__ cand) {
  constexpr int H = KH * 16, NW = H / 32, din = KI * 16, ldh = mlpw_ldh(H);
  extern __shared__ __attribute__((aligned(16))) uint16_t wsm[];
  uint16_t *W3s = wsm;                      // G x ldh
  uint16_t *H1 = wsm + (size_t)G * ldh;     // rows x ldh
  uint16_t *H2 = H1 + MLPW_ROWS * ldh;      // rows x ldh
  float *Zs = reinterpret_cast<float *>(wsm + mlpw_zs_off(H, G));  // rows x (G + 4)
  const int ldz = G + 4;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, col = lane & 31;
  const uint16_t *W1 = Ph + poff[0], *W2 = Ph + poff[2], *W3 = Ph + poff[4];
  const float *b1 = P + poff[1], *b2 = P + poff[3], *b3 = P + poff[5];
  // stationary operands of this wave's neuron tile
  lg_u32x4 wf1[KI], wf2[KH];
  lg_f32x16 bias1, bias2;
#pragma unroll
  for (int ks = 0; ks < KI; ++ks)
    wf1[ks] = *reinterpret_cast<const lg_u32x4 *>(W1 + (long)(w * 32 + col) * din + ks * 16 + 8 * h);
#pragma unroll
  for (int ks = 0; ks < KH; ++ks)
    wf2[ks] = *reinterpret_cast<const lg_u32x4 *>(W2 + (long)(w * 32 + col) * H + ks * 16 + 8 * h);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    bias1[r] = b1[w * 32 + lg_row(r, h)];
    bias2[r] = b2[w * 32 + lg_row(r, h)];
  }
  for (int e = tid; e < G * (H / 8); e += NW * 64) {  // W3 into LDS, 16 bytes at a time
    const int g = e / (H / 8), q = e - g * (H / 8);
    *reinterpret_cast<lg_u32x4 *>(W3s + g * ldh + q * 8) =
        *reinterpret_cast<const lg_u32x4 *>(W3 + (long)g * H + q * 8);
  }
  const long nblk = (rows + MLPW_ROWS - 1) / MLPW_ROWS;
  lg_u32x4 xf[2][KI];
  auto load_x = [&](long blk) {
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const long pr = blk * MLPW_ROWS + rt * 32 + col;
#pragma unroll
      for (int ks = 0; ks < KI; ++ks)
        xf[rt][ks] = (blk < nblk && pr < rows)
                         ? *reinterpret_cast<const lg_u32x4 *>(Xh + pr * din + ks * 16 + 8 * h)
                         : lg_u32x4{0u, 0u, 0u, 0u};
    }
  };
  auto store_relu = [&](uint16_t *dst_rows, int nt, int rt, const lg_f32x16 &acc) {
    uint16_t *dst = dst_rows + (rt * 32 + col) * ldh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float v[4] = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
      mlpf_store4(dst, nt * 32 + 8 * q + 4 * h, v);
    }
  };
  load_x(blockIdx.x);
  __syncthreads();  // W3 in LDS
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    // layer 1
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      lg_f32x16 a = bias1;
#pragma unroll
      for (int ks = 0; ks < KI; ++ks)
        a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(lg_bf16x8, wf1[ks]),
                                                    __builtin_bit_cast(lg_bf16x8, xf[rt][ks]), a, 0, 0, 0);
      store_relu(H1, w, rt, a);
    }
    load_x(blk + gridDim.x);  // next block's rows, in flight through layers 2 and 3
    __syncthreads();
    // layer 2
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      lg_f32x16 a = bias2;
      const uint16_t *src = H1 + (rt * 32 + col) * ldh + 8 * h;
#pragma unroll
      for (int ks = 0; ks < KH; ++ks)
        a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(lg_bf16x8, wf2[ks]),
            __builtin_bit_cast(lg_bf16x8, *reinterpret_cast<const lg_u32x4 *>(src + ks * 16)), a, 0, 0, 0);
      store_relu(H2, w, rt, a);
    }
    __syncthreads();
    // layer 3: (group tile, row tile) pairs
    for (int pr = w; pr < (G / 32) * 2; pr += NW) {
      const int nt = pr >> 1, rt = pr & 1;
      lg_f32x16 a;
#pragma unroll
      for (int r = 0; r < 16; ++r) a[r] = b3[nt * 32 + lg_row(r, h)];
      const uint16_t *wa = W3s + (nt * 32 + col) * ldh + 8 * h;
      const uint16_t *src = H2 + (rt * 32 + col) * ldh + 8 * h;
#pragma unroll
      for (int ks = 0; ks < KH; ++ks)
        a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(lg_bf16x8, *reinterpret_cast<const lg_u32x4 *>(wa + ks * 16)),
            __builtin_bit_cast(lg_bf16x8, *reinterpret_cast<const lg_u32x4 *>(src + ks * 16)), a, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) Zs[(rt * 32 + col) * ldz + nt * 32 + lg_row(r, h)] = a[r];
    }
    __syncthreads();
    // top T per row: NW lanes per row, lane sub scans groups sub, sub+NW, ...
    {
      const int row = tid / NW, sub = tid - row * NW;
      const float *z = Zs + row * ldz;
      float bv[TT];
#pragma unroll
      for (int t = 0; t < TT; ++t) bv[t] = -INFINITY;
      auto insert = [&](float p) {  // sorted descending; packed values are distinct
#pragma unroll
        for (int t = TT - 1; t > 0; --t) bv[t] = __builtin_amdgcn_fmed3f(bv[t - 1], bv[t], p);
        bv[0] = fmaxf(bv[0], p);
      };
      for (int g = sub; g < G; g += NW) {
        float v = z[g];
        v = v == v ? v : -FLT_MAX;  // NaN logits rank last
        const unsigned u = __float_as_uint(v);
        const unsigned id = (u >> 31) ? (unsigned)g : (unsigned)(127 - g);
        insert(__uint_as_float((u & ~127u) | id));
      }
#pragma unroll
      for (int off = 1; off < NW; off <<= 1) {
        float pv[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) pv[t] = __shfl_xor(bv[t], off);
#pragma unroll
        for (int t = 0; t < TT; ++t) insert(pv[t]);
      }
      const long prow = blk * MLPW_ROWS + row;
      if (sub == 0 && prow < rows) {
#pragma unroll
        for (int t = 0; t < TT; ++t) {
          if (t < T) {
            const unsigned u = __float_as_uint(bv[t]), id = u & 127u;
            cand[prow * T + t] = bv[t] == -INFINITY ? -1 : (int)((u >> 31) ? id : 127u - id);
          }
        }
      }
    }
    __syncthreads();  // H1 (and the logits) are rewritten by the next block
  }
}

// Uncertified rows gathered for the reduced-precision full search: FP32
// row, reduced-precision row (stride Dp) and norm bound.
__global__ void k_gather_fb(long cnt, const int *__restrict__ fb, int n, int Dp,
                            const float *__restrict__ X, const uint16_t *__restrict__ Xh,
                            const float2 *__restrict__ xn, float *__restrict__ gX,
                            uint16_t *__restrict__ gXh, float2 *__restrict__ gN) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < cnt * Dp;
       i += (long)gridDim.x * blockDim.x) {
    const long r = i / Dp;
    const int d = (int)(i - r * Dp);
    const long p = fb[r];
    gXh[i] = Xh[p * Dp + d];
    if (d < n) gX[r * n + d] = X[p * n + d];
    if (d == 0) gN[r] = xn[p];
  }
}

__global__ void k_scatter_fb(long cnt, const int *__restrict__ fb, const int *__restrict__ lab,
                             int *__restrict__ labels) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) labels[fb[i]] = lab[i];
}

// ------------------------------------------------------------------ host
static hipError_t lmalloc(void **p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 1); }

#define LCHK(x)                          \
  do {                                   \
    hipError_t _e = (x);                 \
    if (_e != hipSuccess) return _e;     \
  } while (0)

bool learned_active(const Engine *e) {
  const Learned *L = L_of(e);
  return L && e->iters >= L->train_iters;
}

kmeansErrorStatus learned_configure(Engine *e) {
  kmeansOptions &o = e->opt;
  if (o.proposer == KM_PROPOSER_NONE) return KMEANS_SUCCESS;
  if (o.proposer != KM_PROPOSER_MLP && o.proposer != KM_PROPOSER_CENTROID)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (o.groups <= 0) o.groups = std::min(128, e->k);
  if (o.top_groups <= 0) o.top_groups = std::min(4, o.groups);
  if (o.train_iters <= 0) o.train_iters = 3;
  if (o.hidden <= 0) o.hidden = 256;
  if (o.groups > e->k || o.top_groups > 8 || o.top_groups > o.groups || o.tile_rows > 0 ||
      (double)e->m * o.top_groups > 2147483647.0)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (o.proposer == KM_PROPOSER_MLP && (o.groups % 16 || o.hidden % 16))
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (o.proposer == KM_PROPOSER_CENTROID && e->n > 256) return KMEANS_ERROR_INVALID_ARGUMENT;
  Learned *L = new Learned();
  e->learned = L;
  L->proposer = o.proposer;
  L->G = o.groups;
  L->T = o.top_groups;
  L->H = o.hidden;
  L->train_iters = o.train_iters;
  L->din = (e->n + 15) / 16 * 16;
  L->S = std::min<long>(e->m, 262144);
  return KMEANS_SUCCESS;
}

void learned_free(Engine *e) {
  Learned *L = L_of(e);
  if (!L) return;
  void *bufs[] = {L->samp_rows, L->samp_lab, L->samp_tg, L->samp_x, L->gof, L->members, L->gstart,
                  L->gsize, L->gcol, L->permP, L->CtP, L->ncP, L->Gc, L->gnh, L->rad2, L->sep2,
                  L->ncmax, L->P, L->Gr, L->Am, L->Av, L->loss, L->Ph, L->mws, L->Xh, L->H1,
                  L->H2, L->Z, L->cand, L->gcnt, L->ent_pt, L->ent_id, L->elab, L->escore,
                  L->ul_pt, L->ul_best, L->ucnt, L->fb, L->ul_score, L->cnt, L->fbX, L->fbN,
                  L->fbXh, L->fbLab, L->ChP, L->biasP, L->wpre, L->sepTop, L->dpoff};  // goff, gcur, fbcnt
                                                                               // live inside gcnt, ucnt
  for (void *b : bufs) hipFree(b);
  if (L->hcnt) hipHostFree(L->hcnt);
  if (L->hwpre) hipHostFree(L->hwpre);
  delete L;
  e->learned = nullptr;
}

// Sample rows of plain iteration `it` (fixed per iteration, counter-based).
hipError_t learned_collect(Engine *e, int it) {
  Learned *L = L_of(e);
  if (!L || it >= L->train_iters || L->ready) return hipSuccess;
  const int nit = L->train_iters;
  const long ns = L->S * nit;
  if (!L->samp_rows) {
    std::vector<int> rows((size_t)ns);
    for (int t = 0; t < nit; ++t)
      for (long i = 0; i < L->S; ++i)
        rows[(size_t)(i * nit + t)] = (int)(rng_u64(0x1ea7ULL, 8 + (uint64_t)t, (uint64_t)i) % (uint64_t)e->m);
    LCHK(lmalloc((void **)&L->samp_rows, sizeof(int) * ns));
    LCHK(lmalloc((void **)&L->samp_lab, sizeof(int) * ns));
    LCHK(hipMemcpy(L->samp_rows, rows.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    LCHK(hipMemset(L->samp_lab, 0, sizeof(int) * ns));
  }
  hipLaunchKernelGGL(k_collect, dim3(g1(L->S)), dim3(256), 0, e->stream, L->S, nit, it, L->samp_rows,
                     e->dlabels, L->samp_lab);
  return hipGetLastError();
}

// Group membership from a coarse K-means of the centres (the engine's own
// plain FP32 loop, 20 iterations, sampled initial centres).
static kmeansErrorStatus build_groups(Engine *e, Learned *L) {
  const int k = e->k, n = e->n, G = L->G;
  std::vector<int> gof((size_t)k);
  {
    float *Ccopy = nullptr;
    KM_HIP_CHECK(hipMalloc(&Ccopy, sizeof(float) * (size_t)k * n), KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipMemcpyAsync(Ccopy, e->dC, sizeof(float) * (size_t)k * n,
                                hipMemcpyDeviceToDevice, e->stream),
                 KMEANS_ERROR_COMPUTE);
    Engine sub;
    kmeansOptions so{};
    so.precision = KM_PREC_FP32;
    so.device = e->opt.device;
    so.stream = e->stream;
    kmeansErrorStatus st = sub.create(k, n, G, &so);
    if (st == KMEANS_SUCCESS) st = sub.buildData(Ccopy, 1);
    if (st == KMEANS_SUCCESS) st = sub.initialize(0x9a0aULL);
    if (st == KMEANS_SUCCESS) {
      sub.maxIters = 20;
      sub.criterion = -1.0;
      sub.relErr = 1.0;
      st = sub.computeStandard();
    }
    if (st == KMEANS_SUCCESS) st = sub.getLabels(gof.data());
    hipStreamSynchronize(e->stream);
    hipFree(Ccopy);
    if (st != KMEANS_SUCCESS) return st;
  }
  // members sorted by group, ascending centre index within a group
  std::vector<int> gsize(G, 0), gstart(G + 1, 0), members(k), gcol(G), pos(G);
  for (int j = 0; j < k; ++j) gsize[gof[j]]++;
  for (int g = 0; g < G; ++g) gstart[g + 1] = gstart[g] + gsize[g];
  for (int g = 0; g < G; ++g) pos[g] = gstart[g];
  for (int j = 0; j < k; ++j) members[pos[gof[j]]++] = j;
  int col = 0;
  for (int g = 0; g < G; ++g) {
    gcol[g] = col;
    col += (gsize[g] + 3) / 4 * 4;
  }
  L->ldP = col + 256;  // the FP32 kernel reads whole 256-column tiles
  std::vector<int> permP((size_t)L->ldP, -1);
  for (int g = 0; g < G; ++g)
    for (int i = 0; i < gsize[g]; ++i) permP[(size_t)gcol[g] + i] = members[(size_t)gstart[g] + i];
  L->gsize_h = gsize;
  L->gcol_h = gcol;
  auto up = [&](int **dst, const std::vector<int> &src) -> bool {
    if (lmalloc((void **)dst, sizeof(int) * src.size()) != hipSuccess) return false;
    return hipMemcpy(*dst, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up(&L->gof, gof) || !up(&L->members, members) || !up(&L->gstart, gstart) ||
      !up(&L->gsize, gsize) || !up(&L->gcol, gcol) || !up(&L->permP, permP))
    return KMEANS_ERROR_COMPUTE;
  const size_t ldP = (size_t)L->ldP;
  KM_HIP_CHECK(lmalloc((void **)&L->CtP, sizeof(float) * ldP * n), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ncP, sizeof(float) * ldP), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Gc, sizeof(float) * (size_t)G * n), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->gnh, sizeof(float) * G), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->rad2, sizeof(float) * G), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->sep2, sizeof(unsigned) * (size_t)k * G), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->sepTop, sizeof(int2) * (size_t)k * (L->T + 1)), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ncmax, sizeof(float)), KMEANS_ERROR_COMPUTE);
  return KMEANS_SUCCESS;
}

static void init_params(const MlpShape &sh, std::vector<float> &P) {
  size_t off[6];
  mlp_offsets(sh, off);
  const int fan_in[6] = {sh.din, sh.din, sh.h, sh.h, sh.h, sh.h};
  P.resize(mlp_param_count(sh));
  for (int b = 0; b < 6; ++b) {
    const size_t end = b < 5 ? off[b + 1] : P.size();
    const double bound = 1.0 / sqrt((double)fan_in[b]);
    for (size_t i = off[b]; i < end; ++i)
      P[i] = (float)((2.0 * rng_unif(0x31a9ULL, 16 + (uint64_t)b, i) - 1.0) * bound);
  }
}

// MLP training on the sampled rows (Adam, batches of 4096, 3 epochs).
static kmeansErrorStatus train_mlp(Engine *e, Learned *L) {
  const long ns = L->S * L->train_iters;
  const int din = L->din;
  hipStream_t s = e->stream;
  L->sh = MlpShape{din, L->H, L->G};
  L->Pn = mlp_param_count(L->sh);
  const int B = (int)std::min<long>(4096, ns);
  L->mws_bytes = std::max(mlp_ws_bytes(L->sh, B), (size_t)1);
  KM_HIP_CHECK(lmalloc((void **)&L->samp_tg, sizeof(int) * ns), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->samp_x, sizeof(uint16_t) * ns * din), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->P, sizeof(float) * L->Pn), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Gr, sizeof(float) * L->Pn), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Am, sizeof(float) * L->Pn), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Av, sizeof(float) * L->Pn), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Ph, sizeof(uint16_t) * L->Pn), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->loss, sizeof(float) * 4), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc(&L->mws, L->mws_bytes), KMEANS_ERROR_COMPUTE);
  // input scale: 1 / rms coordinate over the sample
  double *dmean = nullptr;
  KM_HIP_CHECK(hipMalloc(&dmean, sizeof(double)), KMEANS_ERROR_COMPUTE);
  hipLaunchKernelGGL(k_mean_q, dim3(1), dim3(256), 0, s, ns, L->samp_rows, e->dq, dmean);
  double mq = 0.0;
  KM_HIP_CHECK(hipMemcpyAsync(&mq, dmean, sizeof(double), hipMemcpyDeviceToHost, s), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_COMPUTE);
  hipFree(dmean);
  const double rms = sqrt(mq / (double)e->n);
  L->xscale = rms > 0.0 && std::isfinite(rms) ? (float)(1.0 / rms) : 1.0f;
  hipLaunchKernelGGL(k_x_bf16, dim3(km_grid(ns * din, 256)), dim3(256), 0, s, ns, e->n, din, e->dX,
                     L->samp_rows, L->xscale, L->samp_x);
  hipLaunchKernelGGL(k_targets, dim3(g1(ns)), dim3(256), 0, s, ns, L->samp_lab, L->gof, L->samp_tg);
  std::vector<float> P0;
  init_params(L->sh, P0);
  KM_HIP_CHECK(hipMemcpyAsync(L->P, P0.data(), sizeof(float) * L->Pn, hipMemcpyHostToDevice, s),
               KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipMemsetAsync(L->Am, 0, sizeof(float) * L->Pn, s), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipMemsetAsync(L->Av, 0, sizeof(float) * L->Pn, s), KMEANS_ERROR_COMPUTE);
  const long nb = ns / B;
  const int epochs = 3;
  const long steps = std::max<long>(1, nb * epochs);
  for (long st = 0; st < steps; ++st) {
    const long b = st % std::max<long>(nb, 1);
    float *lp = st == 0 ? L->loss : (st == steps - 1 ? L->loss + 1 : nullptr);
    const float lr = st < steps / 2 ? 2e-3f : (st < 5 * steps / 6 ? 6e-4f : 2e-4f);
    KM_HIP_CHECK(mlp_fwd_bwd(s, L->sh, B, L->samp_x + (size_t)b * B * din, L->samp_tg + (size_t)b * B,
                             L->P, L->Gr, lp, L->mws),
                 KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(mlp_adam(s, L->Pn, L->P, L->Gr, L->Am, L->Av, lr, (int)st + 1, nullptr),
                 KMEANS_ERROR_COMPUTE);
  }
  if (steps == 1) KM_HIP_CHECK(hipMemcpyAsync(L->loss + 1, L->loss, sizeof(float), hipMemcpyDeviceToDevice, s),
                               KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(params_to_bf16(s, L->Pn, L->P, L->Ph), KMEANS_ERROR_COMPUTE);
  float lh[2] = {0.f, 0.f};
  KM_HIP_CHECK(hipMemcpyAsync(lh, L->loss, sizeof(float) * 2, hipMemcpyDeviceToHost, s), KMEANS_ERROR_COMPUTE);
  // scaled BF16 copy of every point for inference (points do not change)
  KM_HIP_CHECK(lmalloc((void **)&L->Xh, sizeof(uint16_t) * (size_t)e->m * din), KMEANS_ERROR_COMPUTE);
  hipLaunchKernelGGL(k_x_bf16, dim3(km_grid(e->m * din, 256)), dim3(256), 0, s, e->m, e->n, din, e->dX, nullptr,
                     L->xscale, L->Xh);
  {
    size_t off[6];
    mlp_offsets(L->sh, off);
    KM_HIP_CHECK(lmalloc((void **)&L->dpoff, sizeof(off)), KMEANS_ERROR_COMPUTE);
    KM_HIP_CHECK(hipMemcpy(L->dpoff, off, sizeof(off), hipMemcpyHostToDevice), KMEANS_ERROR_COMPUTE);
  }
  L->chunk = std::min<long>(e->m, 1L << 20);
  KM_HIP_CHECK(lmalloc((void **)&L->H1, sizeof(uint16_t) * L->chunk * L->H), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->H2, sizeof(uint16_t) * L->chunk * L->H), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->Z, sizeof(float) * L->chunk * L->G), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipStreamSynchronize(s), KMEANS_ERROR_COMPUTE);
  L->loss_first = lh[0];
  L->loss_last = lh[1];
  return KMEANS_SUCCESS;
}

kmeansErrorStatus learned_prepare(Engine *e) {
  Learned *L = L_of(e);
  if (!L || L->ready) return KMEANS_SUCCESS;
  KM_HIP_CHECK(hipStreamSynchronize(e->stream), KMEANS_ERROR_COMPUTE);
  const auto t0 = std::chrono::steady_clock::now();
  kmeansErrorStatus st = build_groups(e, L);
  if (st != KMEANS_SUCCESS) return st;
  if (L->proposer == KM_PROPOSER_MLP) {
    if (!L->samp_rows) return KMEANS_ERROR_COMPUTE;
    if ((st = train_mlp(e, L)) != KMEANS_SUCCESS) return st;
  }
  const long m = e->m, ne = m * L->T;
  KM_HIP_CHECK(lmalloc((void **)&L->cand, sizeof(int) * ne), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ent_pt, sizeof(int) * ne), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ent_id, sizeof(int) * ne), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->elab, sizeof(int) * ne), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->escore, sizeof(float) * ne), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ul_pt, sizeof(int) * m), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ul_best, sizeof(int) * m), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->ul_score, sizeof(float) * m), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->fb, sizeof(int) * m), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(lmalloc((void **)&L->gcnt, sizeof(int) * L->G * 3), KMEANS_ERROR_COMPUTE);
  L->goff = L->gcnt + L->G;
  L->gcur = L->gcnt + 2 * L->G;
  KM_HIP_CHECK(lmalloc((void **)&L->ucnt, sizeof(int) * 2), KMEANS_ERROR_COMPUTE);
  L->fbcnt = L->ucnt + 1;
  KM_HIP_CHECK(lmalloc((void **)&L->cnt, sizeof(unsigned long long) * 5), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipMemsetAsync(L->cnt, 0, sizeof(unsigned long long) * 5, e->stream), KMEANS_ERROR_COMPUTE);
  KM_HIP_CHECK(hipHostMalloc((void **)&L->hcnt, sizeof(int) * 2 * L->G), KMEANS_ERROR_COMPUTE);
  {
    MfmaView mv;
    L->mfma_groups = e->opt.precision == KM_PREC_BF16 && mfma_view(e, &mv) && !mv.fp16;
    if (L->mfma_groups) {
      KM_HIP_CHECK(lmalloc((void **)&L->ChP, sizeof(uint16_t) * (size_t)L->ldP * mv.Dp),
                   KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(lmalloc((void **)&L->biasP, sizeof(float) * (size_t)L->ldP), KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(lmalloc((void **)&L->wpre, sizeof(int) * (L->G + 1)), KMEANS_ERROR_COMPUTE);
      KM_HIP_CHECK(hipHostMalloc((void **)&L->hwpre, sizeof(int) * (L->G + 1)), KMEANS_ERROR_COMPUTE);
    }
  }
  KM_HIP_CHECK(hipStreamSynchronize(e->stream), KMEANS_ERROR_COMPUTE);
  L->ms_train = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  L->ready = true;
  return KMEANS_SUCCESS;
}

// Per-iteration centre-side data: group-ordered columns, centroids, radii,
// centre separations, max ||c||^2.
static hipError_t refresh_centres(Engine *e, Learned *L) {
  hipStream_t s = e->stream;
  const int k = e->k, n = e->n;
  hipLaunchKernelGGL(k_gather_perm, dim3(g1((long)L->ldP * n)), dim3(256), 0, s, L->ldP, n, L->permP,
                     e->dC, e->dnc, L->CtP, L->ncP);
  hipLaunchKernelGGL(k_group_geom, dim3(L->G), dim3(256), sizeof(float) * n, s, n, L->gstart,
                     L->members, e->dC, L->Gc, L->gnh, L->rad2);
  hipLaunchKernelGGL(k_fill_i32, dim3(g1((long)k * L->G)), dim3(256), 0, s, (long)k * L->G,
                     (int *)L->sep2, (int)0x7f800000);
  dim3 sg((L->ldP + SEP_T - 1) / SEP_T, (k + SEP_T - 1) / SEP_T);
  hipLaunchKernelGGL(k_center_sep, sg, dim3(256), 0, s, k, n, L->G, e->dC, L->CtP, L->ldP, L->permP,
                     L->gof, L->sep2);
  hipLaunchKernelGGL(k_sep_top, dim3(g1(k)), dim3(256), 0, s, k, L->G, L->T + 1, L->sep2, L->sepTop);
  hipLaunchKernelGGL(k_ncmax, dim3(1), dim3(256), 0, s, k, e->dnc, L->ncmax);
  if (L->mfma_groups) {
    MfmaView mv;
    if (!mfma_view(e, &mv)) return hipErrorNotReady;
    const long chunks = (long)L->ldP * (mv.Dp / 8);
    hipLaunchKernelGGL(k_gather_ch, dim3(g1(chunks)), dim3(256), 0, s, L->ldP, mv.Dp, L->permP, mv.Ch,
                       mv.bias, L->ChP, L->biasP);
  }
  return hipGetLastError();
}

static hipError_t propose(Engine *e, Learned *L) {
  hipStream_t s = e->stream;
  const long m = e->m;
  if (L->proposer == KM_PROPOSER_MLP && (L->H == 128 || L->H == 256) && L->G % 32 == 0 &&
      L->G <= 128 && L->T <= 8 && L->din % 16 == 0 && L->din <= 128 && L->din / 16 != 7 &&
      L->din / 16 != 5 && L->din / 16 != 3) {
    // weight-stationary persistent inference (one workgroup per CU)
    const size_t lds = mlpw_lds_bytes(L->H, L->G);
    const dim3 grid((unsigned)std::min<long>((m + MLPW_ROWS - 1) / MLPW_ROWS, (long)num_cus()));
#define MLPW(KH, KI)                                                                              \
  do {                                                                                           \
    if (L->T <= 4) {                                                                             \
      LCHK(hipFuncSetAttribute((const void *)k_mlp_propose_ws<KH, KI, 4>,                        \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));           \
      hipLaunchKernelGGL((k_mlp_propose_ws<KH, KI, 4>), grid, dim3(KH / 2 * 64), lds, s, m, L->G, \
                         L->T, L->Xh, L->Ph, L->P, L->dpoff, L->cand);                           \
    } else {                                                                                     \
      LCHK(hipFuncSetAttribute((const void *)k_mlp_propose_ws<KH, KI, 8>,                        \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));           \
      hipLaunchKernelGGL((k_mlp_propose_ws<KH, KI, 8>), grid, dim3(KH / 2 * 64), lds, s, m, L->G, \
                         L->T, L->Xh, L->Ph, L->P, L->dpoff, L->cand);                           \
    }                                                                                            \
  } while (0)
#define MLPW_KI(KH)                           \
  switch (L->din / 16) {                      \
    case 1: MLPW(KH, 1); break;               \
    case 2: MLPW(KH, 2); break;               \
    case 4: MLPW(KH, 4); break;               \
    case 6: MLPW(KH, 6); break;               \
    default: MLPW(KH, 8); break;              \
  }
    if (L->H == 256) { MLPW_KI(16); } else { MLPW_KI(8); }
#undef MLPW_KI
#undef MLPW
    return hipGetLastError();
  }
  if (L->proposer == KM_PROPOSER_MLP && (L->H == 128 || L->H == 256) && L->G % 32 == 0 &&
      (size_t)MLPF_ROWS * (L->G + 4) * 4 <= (size_t)MLPF_ROWS * (L->H + 8) * 2) {
    // fused inference: no hidden or logit matrices in HBM
    const size_t lds = sizeof(uint16_t) * 2 * MLPF_ROWS * (L->H + 8);
    const dim3 grid((unsigned)((m + MLPF_ROWS - 1) / MLPF_ROWS));
    if (L->H == 256) {
      LCHK(hipFuncSetAttribute((const void *)k_mlp_propose<16>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(k_mlp_propose<16>, grid, dim3(256), lds, s, m, L->din, L->G, L->T, L->Xh, L->Ph,
                         L->P, L->dpoff, L->cand);
    } else {
      hipLaunchKernelGGL(k_mlp_propose<8>, grid, dim3(256), lds, s, m, L->din, L->G, L->T, L->Xh, L->Ph,
                         L->P, L->dpoff, L->cand);
    }
    return hipGetLastError();
  }
  if (L->proposer == KM_PROPOSER_MLP) {
    for (long r0 = 0; r0 < m; r0 += L->chunk) {
      const long R = std::min(L->chunk, m - r0);
      LCHK(mlp_forward(s, L->sh, (int)R, L->Xh + (size_t)r0 * L->din, L->din, L->Ph, L->P, L->H1,
                       L->H2, L->Z));
      hipLaunchKernelGGL(k_top_logits, dim3(g1(R * 16)), dim3(256), 0, s, R, L->G, L->T, L->Z,
                         L->cand + (size_t)r0 * L->T);
    }
    return hipGetLastError();
  }
  const int nv = (e->n + 3) / 4;
  const int NVs[] = {4, 8, 16, 24, 32, 48, 64};
  int NV = 64;
  for (int v : NVs)
    if (v >= nv) { NV = v; break; }
  const int GT = std::min(L->G, 65536 / (16 * NV));
  const size_t lds = sizeof(float) * 4 * (size_t)GT * NV;
#define GT_CASE(V)                                                                             \
  case V:                                                                                      \
    hipFuncSetAttribute((const void *)k_group_top<V>,                                          \
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                 \
    hipLaunchKernelGGL(k_group_top<V>, dim3(g1(m)), dim3(256), lds, s, m, e->n, e->dX, L->G,   \
                       L->Gc, L->gnh, L->gsize, L->T, L->cand);                                \
    break;
  switch (NV) {
    GT_CASE(4) GT_CASE(8) GT_CASE(16) GT_CASE(24) GT_CASE(32) GT_CASE(48) GT_CASE(64)
    default: return hipErrorInvalidValue;
  }
#undef GT_CASE
  return hipGetLastError();
}

hipError_t learned_assign(Engine *e) {
  Learned *L = L_of(e);
  hipStream_t s = e->stream;
  const long m = e->m, ne = m * L->T;
  const int G = L->G;
  LCHK(refresh_centres(e, L));
  LCHK(propose(e, L));
  // entries sorted by group
  hipLaunchKernelGGL(k_fill_i32, dim3(g1(G)), dim3(256), 0, s, (long)G, L->gcnt, 0);
  hipLaunchKernelGGL(k_ent_hist, dim3((unsigned)std::min<long>(g1(ne), 4L * num_cus())), dim3(256),
                     sizeof(int) * G, s, ne, L->cand, G, L->gcnt);
  hipLaunchKernelGGL(k_ent_scan, dim3(1), dim3(64), 0, s, G, L->gcnt, L->goff, L->gcur);
  hipLaunchKernelGGL(k_ent_scatter, dim3((unsigned)((ne + ENT_CHUNK - 1) / ENT_CHUNK)), dim3(256),
                     sizeof(int) * 2 * G, s, ne, L->T, L->cand, G, L->gcur, L->ent_pt, L->ent_id);
  LCHK(hipMemcpyAsync(L->hcnt, L->gcnt, sizeof(int) * 2 * G, hipMemcpyDeviceToHost, s));
  LCHK(hipStreamSynchronize(s));
  for (int g = 0; g < G; ++g)
    if (L->gsize_h[g] > 0) L->pairs += (long long)L->hcnt[g] * L->gsize_h[g];
  if (L->mfma_groups) {
    // BF16 MFMA scores of every named group, certified per entry
    MfmaView mv;
    if (!mfma_view(e, &mv)) return hipErrorNotReady;
    L->hwpre[0] = 0;
    for (int g = 0; g < G; ++g)
      L->hwpre[g + 1] = L->hwpre[g] + (L->gsize_h[g] > 0 ? (L->hcnt[g] + 31) / 32 : 0);
    const int waves = L->hwpre[G];
    LCHK(hipMemcpyAsync(L->wpre, L->hwpre, sizeof(int) * (G + 1), hipMemcpyHostToDevice, s));
    if (waves > 0) {
#define GMF(KS)                                                                                    \
  hipLaunchKernelGGL(k_group_mfma<KS>, dim3((waves + 3) / 4), dim3(256), 0, s, G, L->wpre, L->goff, \
                     L->gcnt, L->gsize, L->gcol, L->permP, L->ent_pt, L->ent_id, mv.Xh, e->dX,      \
                     e->n, L->ChP, L->biasP, e->dC, e->dnc, mv.bnd, mv.xnorm, L->elab, L->escore, \
                     FLT_MAX)
      switch (mv.KS) {
        case 1: GMF(1); break;
        case 2: GMF(2); break;
        case 4: GMF(4); break;
        case 6: GMF(6); break;
        case 8: GMF(8); break;
        default: GMF(16); break;
      }
#undef GMF
      LCHK(hipGetLastError());
    }
  } else {
    // exact canonical scores within each named group
    for (int g = 0; g < G; ++g) {
      const int c = L->hcnt[g], off = L->hcnt[G + g];
      if (c == 0 || L->gsize_h[g] == 0) continue;
      LCHK(launch_assign_f32(s, c, nullptr, e->n, e->dX, L->ent_pt + off, L->gsize_h[g],
                             L->CtP + L->gcol_h[g], L->ldP, L->ncP + L->gcol_h[g], L->elab,
                             L->escore, L->ent_id + off));
    }
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, 2L, L->ucnt, 0);
  BoundParams bp{gamma_n(e->n)};
  hipLaunchKernelGGL(k_combine, dim3(g1(m)), dim3(256), 0, s, m, L->T, L->cand, L->gsize, L->gcol,
                     L->permP, L->elab, L->escore, e->dq, L->sepTop, L->ncmax, bp, e->dlabels,
                     L->ul_pt, L->ul_best, L->ul_score, L->ucnt, L->cnt);
  hipLaunchKernelGGL(k_group_bound, dim3((unsigned)std::min<long>(g1(m), 4L * num_cus())), dim3(256),
                     0, s, L->ucnt, e->n, e->dX, L->T, L->cand, G, L->Gc, L->rad2, L->gsize, e->dq,
                     L->ncmax, bp, L->ul_pt, L->ul_best, L->ul_score, e->dlabels, L->fb, L->fbcnt,
                     L->cnt);
  return hipGetLastError();
}

hipError_t learned_fallback(Engine *e) {
  Learned *L = L_of(e);
  if (e->opt.precision != KM_PREC_FP32 && mfma_ready(e)) {
    // full search of the uncertified rows on the reduced-precision path
    // (itself certified against the FP32 scores, near-ties re-resolved in FP32)
    int cnt = 0;
    LCHK(hipMemcpyAsync(&cnt, L->fbcnt, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    LCHK(hipStreamSynchronize(e->stream));
    const int Dp = mfma_row_stride(e);
    if (cnt > L->fb_cap) {
      const long cap = std::min<long>(e->m, (long)cnt + cnt / 4 + 1024);
      hipFree(L->fbX); hipFree(L->fbN); hipFree(L->fbXh); hipFree(L->fbLab);
      L->fbX = nullptr; L->fbN = nullptr; L->fbXh = nullptr; L->fbLab = nullptr;
      L->fb_cap = 0;
      LCHK(lmalloc((void **)&L->fbX, sizeof(float) * cap * e->n));
      LCHK(lmalloc((void **)&L->fbN, sizeof(float2) * cap));
      LCHK(lmalloc((void **)&L->fbXh, sizeof(uint16_t) * cap * Dp));
      LCHK(lmalloc((void **)&L->fbLab, sizeof(int) * cap));
      L->fb_cap = cap;
    }
    if (cnt > 0) {
      hipLaunchKernelGGL(k_gather_fb, dim3(km_grid((long)cnt * Dp, 256)), dim3(256), 0, e->stream, (long)cnt,
                         L->fb, e->n, Dp, e->dX, mfma_tile_xh(e), mfma_tile_xnorm(e), L->fbX,
                         L->fbXh, L->fbN);
      LCHK(mfma_assign_view(e, L->fbX, L->fbXh, L->fbN, L->fbLab, cnt));
      LCHK(mfma_reresolve_view(e, L->fbX, L->fbXh, L->fbN, L->fbLab, cnt));
      hipLaunchKernelGGL(k_scatter_fb, dim3(g1(cnt)), dim3(256), 0, e->stream, (long)cnt, L->fb,
                         L->fbLab, e->dlabels);
    }
  } else {
    LCHK(launch_assign_f32(e->stream, e->m, L->fbcnt, e->n, e->dX, L->fb, e->k, e->dCt, e->ldc,
                           e->dnc, e->dlabels, nullptr, nullptr, true));
  }
  hipLaunchKernelGGL(k_hits, dim3(g1(e->m)), dim3(256), 0, e->stream, e->m, L->T, L->cand, L->gof,
                     e->dlabels, L->cnt);
  return hipGetLastError();
}

void learned_stats(Engine *e) {
  Learned *L = L_of(e);
  if (!L) return;
  e->stats.ms_train = L->ms_train;
  e->stats.train_loss_first = L->loss_first;
  e->stats.train_loss_last = L->loss_last;
  e->stats.candidate_pairs = L->pairs;
  if (!L->cnt) return;
  unsigned long long c[5] = {0, 0, 0, 0, 0};
  if (hipMemcpyAsync(c, L->cnt, sizeof(c), hipMemcpyDeviceToHost, e->stream) != hipSuccess) return;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return;
  e->stats.certified = (long long)c[0];
  e->stats.certified_group = (long long)c[1];
  e->stats.fallback = (long long)c[2];
  e->stats.proposed = (long long)c[3];
  e->stats.proposal_hits = (long long)c[4];
}

}  // namespace km
