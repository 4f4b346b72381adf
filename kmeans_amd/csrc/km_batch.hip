// Batched independent clusterings: many small K-means problems of uneven N
// in one persistent launch per phase, the whole fixed-count Lloyd loop
// captured as one hipGraph.
//
// No counterpart in the reference (its host class runs one problem per
// object, main/cpp/kmeans.hpp:174-658).  Every instance follows the same
// canonical semantics as the single-problem engine: FP32 fma-chain scores
// (ClosestCenters.cu:60,65-69) with the lowest index among maxima, canonical
// FP64 centre sums (members in ascending point order, chunks of KM_CHUNK
// summed left to right, chunk sums combined the same way), empty clusters
// keep their centre, compactness from the per-centre FP64 moments.  Each
// instance must have fewer than KM_BLOCK points, so its sums are one block.
//
// Layout in HBM: points of all instances concatenated (sum N x d FP32, C
// order), instance i at row xoff[i]; centres instances x k x d; labels and
// FP64 squared norms per point.  Work items of the assign phase are (instance,
// first row) tiles of P points (P chosen by the autotuner); the update phase
// takes one instance per workgroup, largest first.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "km_common.hpp"

namespace km {

#define B_MAXK 1024
#define B_MAXD 64
#define B_MAXP 4096  // largest points per assign work item

struct Batch {
  int nb = 0, d = 0, k = 0;
  kmeansOptions opt{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<long> n_h, xoff_h;
  long total = 0;
  // device
  const float *X = nullptr;
  float *X_owned = nullptr;
  long *xoff = nullptr;
  int *nvec = nullptr;
  float *C = nullptr, *nc = nullptr;
  float2 *bnd = nullptr;  // per instance: max ||c||^2 and max ||c - bf16(c)||^2 (raw FP32 chains)
  int *labels = nullptr;
  // the update's per-instance sort (perm: instance-local indices,
  // label-major), chunk partials and per-cluster counts, offsets and chunk
  // prefixes, in two parities: iteration t writes parity t % 2 and reuses the
  // other's partials where a cluster's members did not change; rstate =
  // [the other parity holds a full update of these points, its parity]
  int *perm[2] = {nullptr, nullptr};
  double *part[2] = {nullptr, nullptr};
  int *gcnt[2] = {nullptr, nullptr}, *goff[2] = {nullptr, nullptr};
  int *gcpre[2] = {nullptr, nullptr};
  int *fd = nullptr, *rstate = nullptr;
  int *imin = nullptr;  // per (instance, cluster): smallest row whose label changed since the last sort
  double *comp = nullptr;
  long *choff = nullptr;     // first chunk slot of each instance
  long slots = 0;            // chunk slots in total
  int4 *desc = nullptr;      // per slot: instance, member range, (local chunk, cluster)
  int *empty = nullptr;
  int *order = nullptr;  // instances, largest first
  int2 *items = nullptr;
  int nitems = 0, P = 0;
  int *counters = nullptr;  // assign, update
  unsigned long long *nexact = nullptr;  // BF16 assign: points searched exactly
  // drift bounds of the BF16 assign: per point (u, l, l3, r), the work
  // items' open lists, per centre drift, per instance the largest drift and
  // largest |x|, the valid flag, the kept / paired / scored counts
  float4 *xb = nullptr;
  uint16_t *olist = nullptr;  // per work item: its open positions (P entries at item * P)
  int *ocnt = nullptr;        // per work item: their count
  size_t ocnt_cap = 0;
  float *drift = nullptr, *dmaxv = nullptr, *xmax = nullptr;
  int *bvalid = nullptr;
  unsigned long long *nkp = nullptr;
  int comp_slots = 0;
  bool data_set = false, centers_set = false;
  // graph of the last run length
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_iters = 0;
  // events recorded in the graph: before the assign, between assign and
  // update, after the update, per iteration
  std::vector<hipEvent_t> gev;
  int iters = 0;
  kmeansBatchStats stats{};
};

static inline unsigned bg1(long n) { return (unsigned)((n + 255) / 256); }

// ------------------------------------------------------------------ kernels
__global__ void k_batch_fill(int *p, int n, int v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// Writes a point's label; when it changes, the point's row (within its
// instance) lowers the smallest changed row of both clusters (im: the
// instance's k entries), from which the update finds where each cluster's
// member list first differs from the previous sort.
__device__ inline void set_label(int *__restrict__ labels, long gi, int lab, int *__restrict__ im,
                                 int row) {
  const int old = labels[gi];
  if (old != lab) {
    atomicMin(im + old, row);
    atomicMin(im + lab, row);
  }
  labels[gi] = lab;
}

// Assign, persistent: each workgroup takes (instance, first row) items from
// a counter until none are left; centres of the current instance are staged
// in LDS ([k][d], float4 when d % 4 == 0) with their -0.5*||c||^2.  Each
// thread scores two points against every centre (canonical chains).
template <int DT>
__global__ __launch_bounds__(256) void k_batch_assign_f32(
    int nitems, const int2 *__restrict__ items, int P, const long *__restrict__ xoff,
    const int *__restrict__ nvec, int d, int k, const float *__restrict__ X,
    const float *__restrict__ C, const float *__restrict__ nc, int *__restrict__ labels,
    int *__restrict__ counter, int *__restrict__ imin) {
  extern __shared__ float sm[];  // k*DT centres, then k biases
  float *Cs = sm, *bs = sm + (size_t)k * DT;
  __shared__ int item_s;
  int cur = -1;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= nitems) break;
    const int2 w = items[it];
    const int inst = w.x;
    if (inst != cur) {
      const float *Ci = C + (size_t)inst * k * d;
      for (int e = threadIdx.x; e < k * DT; e += blockDim.x) {
        const int j = e / DT, q = e - j * DT;
        Cs[e] = q < d ? Ci[(size_t)j * d + q] : 0.0f;
      }
      for (int j = threadIdx.x; j < k; j += blockDim.x) bs[j] = -0.5f * nc[(size_t)inst * k + j];
      cur = inst;
    }
    __syncthreads();
    const long n = nvec[inst];
    const long rows = min((long)P, n - w.y);
    const long base = xoff[inst] + w.y;
    for (long p0 = 2 * threadIdx.x; p0 < rows; p0 += 2 * blockDim.x) {
      const bool two = p0 + 1 < rows;
      float x0[DT], x1[DT];
#pragma unroll
      for (int q = 0; q < DT; ++q) {
        x0[q] = q < d ? X[(base + p0) * d + q] : 0.0f;
        x1[q] = (q < d && two) ? X[(base + p0 + 1) * d + q] : 0.0f;
      }
      float b0 = -INFINITY, b1 = -INFINITY;
      int i0 = 0, i1 = 0;
      for (int j = 0; j < k; ++j) {
        float s0 = bs[j], s1 = s0;
        const float *c = Cs + (size_t)j * DT;
#pragma unroll
        for (int q = 0; q < DT; ++q) {
          if (q < d) {
            const float cv = c[q];
            s0 = __builtin_fmaf(x0[q], cv, s0);
            s1 = __builtin_fmaf(x1[q], cv, s1);
          }
        }
        if (s0 > b0) { b0 = s0; i0 = j; }
        if (s1 > b1) { b1 = s1; i1 = j; }
      }
      int *im = imin + (size_t)inst * k;
      set_label(labels, base + p0, i0, im, (int)(w.y + p0));
      if (two) set_label(labels, base + p0 + 1, i1, im, (int)(w.y + p0 + 1));
    }
    __syncthreads();  // item_s and the staged centres are reused
  }
}

// Assign for d == DT (a multiple of 4): centres in LDS as float4 rows,
// PPT points per thread (rows tid + 256*u of the item), every score one fma
// chain in ascending d.
template <int DT, int PPT>
__global__ __launch_bounds__(256) void k_batch_assign_f32x(
    int nitems, const int2 *__restrict__ items, int P, const long *__restrict__ xoff,
    const int *__restrict__ nvec, int k, const float *__restrict__ X,
    const float *__restrict__ C, const float *__restrict__ nc, int *__restrict__ labels,
    int *__restrict__ counter, int *__restrict__ imin) {
  constexpr int V = DT / 4;
  extern __shared__ float4 sm4[];  // k*V centre float4s, then k biases
  float4 *Cs = sm4;
  float *bs = reinterpret_cast<float *>(sm4 + (size_t)k * V);
  __shared__ int item_s;
  int cur = -1;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= nitems) break;
    const int2 w = items[it];
    const int inst = w.x;
    if (inst != cur) {
      const float4 *Ci = reinterpret_cast<const float4 *>(C + (size_t)inst * k * DT);
      for (int e = threadIdx.x; e < k * V; e += blockDim.x) Cs[e] = Ci[e];
      for (int j = threadIdx.x; j < k; j += blockDim.x) bs[j] = -0.5f * nc[(size_t)inst * k + j];
      cur = inst;
    }
    __syncthreads();
    const long rows = min((long)P, (long)nvec[inst] - w.y);
    const long base = xoff[inst] + w.y;
    for (long r0 = 0; r0 < rows; r0 += 256 * PPT) {
      float4 x[PPT][V];
      bool ok[PPT];
#pragma unroll
      for (int u = 0; u < PPT; ++u) {
        const long p = r0 + threadIdx.x + 256 * u;
        ok[u] = p < rows;
        const float4 *xr = reinterpret_cast<const float4 *>(X + (base + (ok[u] ? p : 0)) * DT);
#pragma unroll
        for (int v = 0; v < V; ++v) x[u][v] = xr[v];
      }
      float bv[PPT];
      int bi[PPT];
#pragma unroll
      for (int u = 0; u < PPT; ++u) { bv[u] = -INFINITY; bi[u] = 0; }
      for (int j = 0; j < k; ++j) {
        float sc[PPT];
        const float b0 = bs[j];
#pragma unroll
        for (int u = 0; u < PPT; ++u) sc[u] = b0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const float4 c = Cs[(size_t)j * V + v];
#pragma unroll
          for (int u = 0; u < PPT; ++u) {
            sc[u] = __builtin_fmaf(x[u][v].x, c.x, sc[u]);
            sc[u] = __builtin_fmaf(x[u][v].y, c.y, sc[u]);
            sc[u] = __builtin_fmaf(x[u][v].z, c.z, sc[u]);
            sc[u] = __builtin_fmaf(x[u][v].w, c.w, sc[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < PPT; ++u)
          if (sc[u] > bv[u]) { bv[u] = sc[u]; bi[u] = j; }
      }
      int *im = imin + (size_t)inst * k;
#pragma unroll
      for (int u = 0; u < PPT; ++u)
        if (ok[u])
          set_label(labels, base + r0 + threadIdx.x + 256 * u, bi[u], im,
                    (int)(w.y + r0 + threadIdx.x + 256 * u));
    }
    __syncthreads();
  }
}

// Assign on BF16 MFMA for d <= 32 (KS = 1 or 2 k-steps of 16), labels equal
// to the FP32 labels.  Persistent over (instance, first row) items like the
// FP32 kernel; per instance the LDS holds the BF16 centre rows (k padded to
// 32, zero-padded dimensions), the biases -0.5*nc (-inf past k), the FP32
// centres and norms, and the bound constants max ||c|| and max nc.  Each
// wave scores 32 points at a time with v_mfma_f32_32x32x16_bf16 (points
// converted to BF16 on the fly), keeps a top-3, computes the exact FP32
// chains of the first two (lane halves), and labels the point when the
// better one (lowest index on ties) exceeds the third approximate score + E,
// E being the BF16 forward-error bound of the single-problem certificate.
// Points that fail are searched exactly by the whole wave (64 lanes over the
// k centres, canonical chains, argmax with the tie rule).
typedef __attribute__((ext_vector_type(8))) __bf16 b_bf16x8;
typedef __attribute__((ext_vector_type(16))) float b_f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int b_u32x4;
__device__ inline int b_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }
__device__ inline uint16_t b_bf16(float f) { return bf16_rne(f); }

// An instance's certificate constants from its raw FP32 chains (bnd = max
// ||c||^2, max ||c - bf16(c)||^2): max ||c||^2 and the bounds of ||c|| and
// ||c - bf16(c)||, each chain within gamma_d of its sum.
__device__ inline void inst_bounds(float2 bb, int d, float *cmax, float *ncmax, float *dcmax) {
  const float nm = bb.x * (1.0f + (float)d * 1.2e-7f + 1e-6f) + 1e-30f;
  *ncmax = nm;
  *cmax = sqrtf(nm) * (1.0f + 1e-6f);
  *dcmax = sqrtf(bb.y * (1.0f + (float)d * 1.2e-7f + 1e-6f)) * (1.0f + 1e-6f) + 1e-30f;
}

// Drift-bound classification of one work item's points, one workgroup per
// item ahead of the BF16 assign (a kernel of its own, at full occupancy, so
// that its gathers overlap across many items): kept points get their moved
// bounds; two-candidate points are decided by the FP32 chains of their label
// and runner-up (centres and norms staged in LDS, the values the assign
// uses), lowest index on equal chains, and kept without them when neither
// centre moved; the open ones go to the item's list (olist + item * P, count
// ocnt[item]), which the assign then scores.  The two-candidate points of a
// pass are listed in LDS and decided after it by consecutive threads, so
// their row loads are issued together rather than one divergent step at a
// time.
// Nothing to do before the first assign of the current data and centres
// (*bvalid == 0): the assign then scores every point.
#define CL_PPT 4
__global__ __launch_bounds__(256) void k_batch_classify(
    int nitems, const int2 *__restrict__ items, int P, const long *__restrict__ xoff,
    const int *__restrict__ nvec, int d, int k, const float *__restrict__ X,
    const float *__restrict__ C, const float *__restrict__ nc, const float2 *__restrict__ bnd,
    int *__restrict__ labels, float4 *__restrict__ xb, const int *__restrict__ bvalid,
    const float *__restrict__ drift, const float *__restrict__ dmaxv,
    const float *__restrict__ xmax, unsigned long long *__restrict__ nkp, int *__restrict__ imin,
    uint16_t *__restrict__ olist, int *__restrict__ ocnt) {
  extern __shared__ float csm[];  // the instance's centres (k x cs), norms and drifts (k each)
  __shared__ int cnt_s, pc_s[2];
  __shared__ int pl[256 * CL_PPT];  // the pass's two-candidate points
  __shared__ unsigned kp_s[2];
  const int it = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  if (it >= nitems || !*bvalid) return;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (tid == 0) {
    cnt_s = 0;
    pc_s[0] = pc_s[1] = 0;
    kp_s[0] = kp_s[1] = 0;
  }
  __syncthreads();
  const int2 w = items[it];
  const int inst = w.x;
  const long rows = min((long)P, (long)nvec[inst] - w.y);
  const long base = xoff[inst] + w.y;
  const float Df = (float)d;
  float cmax, ncmax, dcmax;
  inst_bounds(bnd[inst], d, &cmax, &ncmax, &dcmax);
  // centres in LDS at an odd row stride (the two-candidate chains read rows
  // of different centres across the lanes), their norms and drifts
  const int cs = d | 1;
  float *Cs = csm, *ncs = csm + (size_t)k * cs, *dr = ncs + k;
  {
    const float *Ci = C + (size_t)inst * k * d;
    for (int e = tid; e < k * d; e += blockDim.x) {
      const int j = e / d;
      Cs[(size_t)j * cs + (e - j * d)] = Ci[e];
    }
    for (int j = tid; j < k; j += blockDim.x) {
      ncs[j] = nc[(size_t)inst * k + j];
      dr[j] = drift[(size_t)inst * k + j];
    }
  }
  __syncthreads();
  const float dm = dmaxv[inst];
  const float e32i = Df * 5.9604645e-8f * 1.0001f * (0.5f * ncmax + xmax[inst] * cmax);
  const float t4 = 4.0f * e32i * 1.0001f + 1e-30f;
  uint16_t *ol = olist + (size_t)it * P;
  unsigned nk = 0, np = 0;
  // CL_PPT points per thread per pass: their labels and bounds are loaded
  // together
  const bool vec = (d & 3) == 0;
  int pp = 0;  // pass parity: the pair count of this pass is pc_s[pp]
  for (long q0 = 0; q0 < rows; q0 += 256 * CL_PPT, pp ^= 1) {
    int la[CL_PPT];
    float4 bb[CL_PPT];
#pragma unroll
    for (int u = 0; u < CL_PPT; ++u) {
      const long q = q0 + 256 * u + tid;
      la[u] = 0;
      bb[u] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
      if (q < rows) {
        la[u] = labels[base + q];
        bb[u] = xb[base + q];
      }
    }
    float da[CL_PPT], drr[CL_PPT];
#pragma unroll
    for (int u = 0; u < CL_PPT; ++u) {
      const int r = __float_as_int(bb[u].w);
      da[u] = dr[la[u]];
      drr[u] = (r >= 0 && r < k) ? dr[r] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < CL_PPT; ++u) {
      // the class computed for every lane (rows past the item carry r = -1
      // and zero bounds) and masked by the row's validity, without branches
      const long q = q0 + 256 * u + tid;
      const bool valid = q < rows;
      const int a = la[u];
      const float4 b4 = bb[u];
      const int r = __float_as_int(b4.w);
      const float uu = (b4.x + da[u]) * (1.0f + 1e-6f);
      const float l = fmaxf((b4.y - dm) * (1.0f - 1e-6f), 0.0f);
      const float l3 = fmaxf((b4.z - dm) * (1.0f - 1e-6f), 0.0f);
      const bool keep1 = l > uu && (l - uu) * (l + uu) * (1.0f - 1e-6f) > t4;
      const bool two = r >= 0 && r < k && r != a && l3 > uu &&
                       (l3 - uu) * (l3 + uu) * (1.0f - 1e-6f) > t4;
      const bool still = da[u] == 0.0f && drr[u] == 0.0f;
      const int cls = !valid ? 3 : keep1 ? 0 : two ? (still ? 0 : 1) : 2;
      if (cls == 0) {
        xb[base + q] = make_float4(uu, l, l3, b4.w);
        ++nk;
      }
      const bool open = cls == 2;
      const uint64_t mk = __ballot(open);
      if (mk) {
        const int lead = __ffsll((long long)mk) - 1;
        int ob = 0;
        if (lane == lead) ob = atomicAdd(&cnt_s, __popcll(mk));
        ob = __shfl(ob, lead);
        if (open) ol[ob + __popcll(mk & below)] = (uint16_t)q;
      }
      const bool pair = cls == 1;
      const uint64_t pk = __ballot(pair);
      if (pk) {
        const int lead = __ffsll((long long)pk) - 1;
        int ob = 0;
        if (lane == lead) ob = atomicAdd(&pc_s[pp], __popcll(pk));
        ob = __shfl(ob, lead);
        if (pair) pl[ob + __popcll(pk & below)] = (int)q;
      }
    }
    __syncthreads();
    // the pass's two-candidate points: the plain path's chains of the label
    // a and the runner-up r, and the point's own norm bound from the same
    // row (label and bounds read again; nothing above wrote them)
    const int npr = pc_s[pp];
    if (tid == 0) pc_s[pp ^ 1] = 0;  // last read in the previous pass
    for (int e = tid; e < npr; e += 256) {
      const int q = pl[e];
      const long gp = base + q;
      const int a = labels[gp];
      const float4 b4 = xb[gp];
      const int r = __float_as_int(b4.w);
      const float l3 = fmaxf((b4.z - dm) * (1.0f - 1e-6f), 0.0f);
      const float *xr = X + gp * d;
      const float *ca = Cs + (size_t)a * cs, *cb = Cs + (size_t)r * cs;
      float sa = -0.5f * ncs[a], sr = -0.5f * ncs[r], sq = 0.0f;
      for (int dq = 0; dq < d; dq += 16) {
        float xv[16];
        if (vec) {
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const float4 t4v = dq + 4 * g < d ? *reinterpret_cast<const float4 *>(xr + dq + 4 * g)
                                              : make_float4(0.f, 0.f, 0.f, 0.f);
            xv[4 * g] = t4v.x; xv[4 * g + 1] = t4v.y; xv[4 * g + 2] = t4v.z; xv[4 * g + 3] = t4v.w;
          }
        } else {
#pragma unroll
          for (int v = 0; v < 16; ++v) xv[v] = dq + v < d ? xr[dq + v] : 0.0f;
        }
        if (dq + 16 <= d) {
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            sa = __builtin_fmaf(xv[v], ca[dq + v], sa);
            sr = __builtin_fmaf(xv[v], cb[dq + v], sr);
            sq = __builtin_fmaf(xv[v], xv[v], sq);
          }
        } else {
#pragma unroll
          for (int v = 0; v < 16; ++v)
            if (dq + v < d) {
              sa = __builtin_fmaf(xv[v], ca[dq + v], sa);
              sr = __builtin_fmaf(xv[v], cb[dq + v], sr);
              sq = __builtin_fmaf(xv[v], xv[v], sq);
            }
        }
      }
      const bool tr = sr > sa || (sr == sa && r < a);
      const int L = tr ? r : a, O = tr ? a : r;
      const float sL = tr ? sr : sa, sO = tr ? sa : sr;
      if (L != a) {
        atomicMin(imin + (size_t)inst * k + a, (int)(w.y + q));
        atomicMin(imin + (size_t)inst * k + L, (int)(w.y + q));
      }
      labels[gp] = L;
      const float xn = sqrtf(sq * (1.0f + Df * 1.2e-7f + 1e-6f)) * (1.0f + 1e-6f) + 1e-30f;
      xb[gp] = make_float4(dist_upper(xn, sL - e32i), fminf(dist_lower(xn, sO + e32i, Df), l3),
                           l3, __int_as_float(O));
      ++np;
    }
    __syncthreads();
  }
  if (nk) atomicAdd(&kp_s[0], nk);
  if (np) atomicAdd(&kp_s[1], np);
  __syncthreads();
  if (tid == 0) {
    ocnt[it] = cnt_s;
    if (kp_s[0]) atomicAdd(nkp, (unsigned long long)kp_s[0]);
    if (kp_s[1]) atomicAdd(nkp + 1, (unsigned long long)kp_s[1]);
  }
}

//
// Drift bounds across iterations (xb non-null, used when *bvalid): per point
// (u, l, l3, r) as in the single-problem engine (u >= |x - c_label|, l <= the
// distance to every other centre, r the runner-up, l3 <= the distance to
// every centre but the label and r), valid for the centres of the last
// assign; k_batch_finalize leaves each centre's drift and each instance's
// largest.  k_batch_classify has sorted each work item's points into kept
// (the moved bounds separate every other centre by more than the FP32
// chains' error), two-candidate (decided there) and open, whose list
// (olist + item * P, ocnt) is all this kernel scores.  E32 uses the
// instance's largest |x| (xmax).  Counts of kept and paired points go to
// nkp[0..1] (from the classification), of scored points to nkp[2].
// (four waves per SIMD for one k-step: 128 VGPRs, a few spilled outside the
// tile loop; two k-steps keep the compiler's choice)
template <int KS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KS == 1 ? 4 : 1)))
void k_batch_assign_bf16(
    int nitems, const int2 *__restrict__ items, int P, const long *__restrict__ xoff,
    const int *__restrict__ nvec, int d, int k, const float *__restrict__ X,
    const float *__restrict__ C, const float *__restrict__ nc, const float2 *__restrict__ bnd,
    int *__restrict__ labels, int *__restrict__ counter, unsigned long long *__restrict__ nexact,
    float big, float4 *__restrict__ xb, const int *__restrict__ bvalid,
    const float *__restrict__ drift, const float *__restrict__ dmaxv,
    const float *__restrict__ xmax, unsigned long long *__restrict__ nkp, int *__restrict__ imin,
    const uint16_t *__restrict__ olist, const int *__restrict__ ocnt) {
  constexpr int Dp = KS * 16;
  const int Kp = (k + 31) / 32 * 32;
  extern __shared__ __attribute__((aligned(16))) unsigned char bsm[];
  // LDS rows are laid out so that the lanes of a wave read distinct banks:
  // the BF16 chunks as Dp/8 planes of Kp 16-byte entries (lane col of half h
  // reads plane 2 ks + h, entry t * 32 + col), the FP32 centres at an odd
  // row stride cs (lanes reading different centres' dimension q)
  const int cs = d | 1;
  b_u32x4 *Ch = reinterpret_cast<b_u32x4 *>(bsm);                 // (Dp/8) x Kp chunks
  float *bias = reinterpret_cast<float *>(Ch + (size_t)Kp * (Dp / 8));  // Kp
  float *Cf = bias + Kp;                                            // k x cs
  float *ncs = Cf + (size_t)k * cs;                                 // k
  __shared__ int item_s;
  __shared__ float bndv[3];
  __shared__ unsigned long long exact_ct, kp_ct;
  const int tid = threadIdx.x, lane = tid & 63, h = lane >> 5, col = lane & 31;
  const bool bv = xb && *bvalid;
  if (tid == 0) {
    exact_ct = 0;
    kp_ct = 0;
  }
  int cur = -1;
  for (;;) {
    if (tid == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= nitems) break;
    const int2 w = items[it];
    const int inst = w.x;
    if (inst != cur) {
      // each centre value is read once (two 16-byte loads per 8-value chunk
      // when d is a multiple of 8) for both its BF16 chunk and its FP32 row
      const float *Ci = C + (size_t)inst * k * d;
      const bool v8 = (d & 7) == 0;
      for (int e = tid; e < Kp * (Dp / 8); e += 256) {
        const int j = e / (Dp / 8), q8 = (e - j * (Dp / 8)) * 8;
        float v[8];
        if (v8) {
          float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
          if (j < k && q8 < d) {
            lo = *reinterpret_cast<const float4 *>(Ci + (size_t)j * d + q8);
            hi = *reinterpret_cast<const float4 *>(Ci + (size_t)j * d + q8 + 4);
          }
          v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = lo.w;
          v[4] = hi.x; v[5] = hi.y; v[6] = hi.z; v[7] = hi.w;
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = (j < k && q8 + u < d) ? Ci[(size_t)j * d + q8 + u] : 0.0f;
        }
        Ch[(size_t)(q8 / 8) * Kp + j] = b_u32x4{bf16x2_rne(v[0], v[1]), bf16x2_rne(v[2], v[3]),
                                                bf16x2_rne(v[4], v[5]), bf16x2_rne(v[6], v[7])};
        if (j < k)
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (q8 + u < d) Cf[(size_t)j * cs + q8 + u] = v[u];
      }
      for (int j = tid; j < Kp; j += 256) {
        const float n2 = j < k ? nc[(size_t)inst * k + j] : 0.0f;
        bias[j] = j < k ? -0.5f * n2 : -FLT_MAX;
        if (j < k) ncs[j] = n2;
      }
      // the instance's bounds (k_batch_bounds / k_batch_finalize): the nc
      // chains are within gamma_d of the sums, and so are the chains of the
      // BF16 residuals c - bf16(c) (exact in FP32)
      if (tid == 0) inst_bounds(bnd[inst], d, &bndv[0], &bndv[1], &bndv[2]);
      cur = inst;
    }
    __syncthreads();
    const float cmax = bndv[0], ncmax = bndv[1], dcmax = bndv[2];
    const long rows = min((long)P, (long)nvec[inst] - w.y);
    const long base = xoff[inst] + w.y;
    const int wv0 = tid >> 6;
    const float Df = (float)d;
    // the item's positions to score: the open list when the bounds hold
    const long nopen = bv ? (long)ocnt[it] : rows;
    const uint16_t *ol = olist + (size_t)it * P;
    // item position -> row within the item
    auto row_of = [&](long pos) -> long { return bv ? (long)ol[pos] : pos; };
    // this half's 8 dimensions per k-step of the point in row pn (zero past
    // d or past the rows); the next 32 points' values are loaded while the
    // current ones are scored
    const bool vec = (d & 3) == 0;  // rows 16-byte aligned: float4 loads
    auto load_half = [&](long pn, float *v) {
      const bool okn = pn < nopen;
      const float *xn = X + (base + (okn ? row_of(pn) : 0)) * d;
      if (vec) {
#pragma unroll
        for (int g = 0; g < 2 * KS; ++g) {
          const int d0 = (g >> 1) * 16 + 8 * h + 4 * (g & 1);
          const float4 t4 = (okn && d0 < d) ? *reinterpret_cast<const float4 *>(xn + d0)
                                            : make_float4(0.f, 0.f, 0.f, 0.f);
          v[4 * g] = t4.x; v[4 * g + 1] = t4.y; v[4 * g + 2] = t4.z; v[4 * g + 3] = t4.w;
        }
        return;
      }
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int d0 = ks * 16 + 8 * h + u;
          v[ks * 8 + u] = (okn && d0 < d) ? xn[d0] : 0.0f;
        }
    };
    float nxt[KS * 8];
    load_half(wv0 * 32 + col, nxt);
    for (long p0 = wv0 * 32; p0 < nopen; p0 += 4 * 32) {
      const long pos = p0 + col;
      const bool ok = pos < nopen;
      const long pr = ok ? row_of(pos) : 0;  // the point's row within the item
      const float *xr = X + (base + pr) * d;
      float cur_x[KS * 8];
#pragma unroll
      for (int i = 0; i < KS * 8; ++i) cur_x[i] = nxt[i];
      if (p0 + 4 * 32 < nopen) load_half(p0 + 4 * 32 + col, nxt);
      // the point's BF16 fragments (this half's 8 dimensions per k-step) and
      // its ||x||^2 part
      b_u32x4 bf[KS];
      float sq = 0.0f, se = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        uint32_t wv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float v0 = cur_x[ks * 8 + 2 * u];
          const float v1 = cur_x[ks * 8 + 2 * u + 1];
          sq = __builtin_fmaf(v0, v0, sq);
          sq = __builtin_fmaf(v1, v1, sq);
          const uint32_t hw = bf16x2_rne(v0, v1);  // one conversion for both
          const float r0 = v0 - __uint_as_float(hw << 16);
          const float r1 = v1 - __uint_as_float(hw & 0xFFFF0000u);
          se = __builtin_fmaf(r0, r0, se);
          se = __builtin_fmaf(r1, r1, se);
          wv[u] = hw;
        }
        bf[ks] = b_u32x4{wv[0], wv[1], wv[2], wv[3]};
      }
      sq += __shfl_xor(sq, 32);
      se += __shfl_xor(se, 32);
      // upper bounds of |x| and |x - bf16(x)|: both halves' chains within
      // gamma_d of the sums
      const float inf = 1.0f + (float)d * 1.2e-7f + 1e-6f;
      const float xn = sqrtf(sq * inf) * (1.0f + 1e-6f) + 1e-30f;
      const float ex = sqrtf(se * inf) * (1.0f + 1e-6f) + 1e-30f;
      // top 3 of the packed scores (register id in the low 4 bits, three
      // v_med3 per score, max as med3 with a run-time +FLT_MAX) and the tiles
      // of the first two.  Centres past k have bias -FLT_MAX and zero BF16
      // rows, so their scores are finite (-FLT_MAX) and pack without a clamp,
      // below every real score
      float b1 = -FLT_MAX, b2 = -FLT_MAX, b3 = -FLT_MAX;
      int t1 = 0, t2 = 0;
      auto score_tile = [&](int t) {
        b_f32x16 acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // rows 8q + 4h .. +3: one 16-byte read
          const float4 b4 = *reinterpret_cast<const float4 *>(bias + t * 32 + 8 * q + 4 * h);
          acc[4 * q] = b4.x; acc[4 * q + 1] = b4.y; acc[4 * q + 2] = b4.z; acc[4 * q + 3] = b4.w;
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const b_u32x4 a = Ch[(size_t)(ks * 2 + h) * Kp + t * 32 + col];
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b_bf16x8, a),
                                                         __builtin_bit_cast(b_bf16x8, bf[ks]), acc, 0, 0, 0);
        }
        return acc;
      };
      auto epilogue = [&](int t, const b_f32x16 &acc) {
        const float o1 = b1, o2 = b2;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float pv = __uint_as_float((__float_as_uint(acc[r]) & 0xFFFFFFF0u) | (unsigned)r);
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
      // the matrix core scores tile t+1 while the VALU runs tile t's
      // epilogue; the last one or two tiles are peeled so that no tile past
      // the end is scored
      const int nT = Kp / 32;
      b_f32x16 A = score_tile(0), B;
      int t = 0;
      for (; t + 2 < nT; t += 2) {
        B = score_tile(t + 1);
        epilogue(t, A);
        A = score_tile(t + 2);
        epilogue(t + 1, B);
      }
      if (t + 1 < nT) {  // t = nT - 2
        B = score_tile(t + 1);
        epilogue(t, A);
        epilogue(t + 1, B);
      } else {  // t = nT - 1
        epilogue(t, A);
      }
      const bool real1 = b1 > -FLT_MAX * 0.5f, real2 = b2 > -FLT_MAX * 0.5f;
      int i1 = real1 ? t1 * 32 + b_row((int)(__float_as_uint(b1) & 15u), h) : -1;
      int i2 = real2 ? t2 * 32 + b_row((int)(__float_as_uint(b2) & 15u), h) : -1;
      // bitwise-equal packed values make the tile tracking ambiguous: such
      // points go to the exact search
      bool tie = (__float_as_uint(b1) == __float_as_uint(b2)) ||
                 (__float_as_uint(b2) == __float_as_uint(b3));
      tie = tie || (__shfl_xor((int)tie, 32) != 0);
      // merge halves; half 0's result is used by both
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
        const bool mt = (__float_as_uint(m1) == __float_as_uint(m2)) ||
                        (__float_as_uint(m2) == __float_as_uint(m3));
        b1 = __shfl(m1, col);
        b2 = __shfl(m2, col);
        b3 = __shfl(m3, col);
        i1 = __shfl(k1, col);
        i2 = __shfl(k2, col);
        tie = tie || (__shfl((int)mt, col) != 0);
      }
      // the point's FP32 row in registers (zero past d), loaded only now so
      // that it is not held across the tile loop
      float xv[Dp];
#pragma unroll
      for (int q4 = 0; q4 < Dp / 4; ++q4) {
        float4 t4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok && 4 * q4 < d) {
          if (vec) {
            t4 = reinterpret_cast<const float4 *>(xr)[q4];
          } else {
            t4.x = xr[4 * q4];
            if (4 * q4 + 1 < d) t4.y = xr[4 * q4 + 1];
            if (4 * q4 + 2 < d) t4.z = xr[4 * q4 + 2];
            if (4 * q4 + 3 < d) t4.w = xr[4 * q4 + 3];
          }
        }
        xv[4 * q4] = t4.x; xv[4 * q4 + 1] = t4.y; xv[4 * q4 + 2] = t4.z; xv[4 * q4 + 3] = t4.w;
      }
      // exact chains of the two candidates
      const int cj = h ? i2 : i1;
      float sx = -INFINITY;
      if (ok && cj >= 0 && cj < k) {
        float sc = -0.5f * ncs[cj];
        const float *cr = Cf + (size_t)cj * cs;
        if (d == Dp) {
#pragma unroll
          for (int q = 0; q < Dp; ++q) sc = __builtin_fmaf(xv[q], cr[q], sc);
        } else {
#pragma unroll
          for (int q = 0; q < Dp; ++q)
            if (q < d) sc = __builtin_fmaf(xv[q], cr[q], sc);
        }
        sx = sc;
      }
      const float sy = __shfl_xor(sx, 32);
      const int jy = __shfl_xor(cj, 32);
      bool fail = false;
      if (ok && h == 0) {
        int lab = cj, oth = jy;
        float sb = sx, so = sy;
        if (jy >= 0 && jy < k && (sy > sb || (sy == sb && jy < lab))) {
          sb = sy;
          lab = jy;
          oth = cj;
          so = sx;
        }
        // forward-error bound of the BF16 score (|x.c - xh.ch| <= |dx| |c| +
        // |xh| |dc| by Cauchy-Schwarz, dx and dc the rounding errors), the
        // matrix core's accumulation and the FP32 chain
        const float g32 = Df * 5.9604645e-8f * 1.0001f;
        const float gmf = (Df + 1.0f) * 2.3841858e-7f * 1.001f;
        const float rnd = ex * cmax + (xn + ex) * dcmax;
        const float mfe = gmf * (0.5f * ncmax + (xn + ex) * (cmax + dcmax));
        const float fpe = g32 * (0.5f * ncmax + xn * cmax);
        const float E = (rnd + mfe + fpe) * 1.001f + 1e-30f;
        const bool has3 = b3 > -FLT_MAX * 0.5f;
        // packed low bits (< 16 ulp) and FP32 rounding of the comparison
        const float slack = (fabsf(b1) + fabsf(b2) + (has3 ? fabsf(b3) : 0.0f)) * 1.6e-5f;
        const float thr = has3 ? b3 + E + slack : -INFINITY;
        if (!tie && lab >= 0 && sb > -INFINITY && sb > thr)
          set_label(labels, base + pr, lab, imin + (size_t)inst * k, (int)(w.y + pr));
        else
          fail = true;
        if (xb) {
          // fresh bounds as in the single-problem kernel: the label's chain
          // minus E, the other candidate's chain and the third approximate
          // score plus E (and the slack); a point searched exactly below gets
          // bounds that never hold
          float4 nb4 = make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
          if (!fail) {
            const float hi3 = has3 ? b3 + E + slack : -INFINITY;
            const float hi = fmaxf(hi3, so + E);
            const int rr = (oth >= 0 && oth < k && oth != lab) ? oth : -1;
            nb4 = make_float4(dist_upper(xn, sb - E), dist_lower(xn, hi, Df),
                              hi3 == -INFINITY ? INFINITY : dist_lower(xn, hi3, Df),
                              __int_as_float(rr));
          }
          xb[base + pr] = nb4;
        }
      }
      if (bv || xb) {  // scored points, one count per wave
        const uint64_t sm = __ballot(ok && h == 0);
        if (lane == 0 && sm) atomicAdd(&kp_ct, (unsigned long long)__popcll(sm));
      }
      // exact search of the points that failed, by the whole wave
      uint64_t fm = __ballot(fail);
      if (lane == 0 && fm) atomicAdd(&exact_ct, (unsigned long long)__popcll(fm));
      while (fm) {
        const int src = __ffsll((long long)fm) - 1;
        fm &= fm - 1;
        const long pp = row_of(p0 + (src & 31));
        float xq[Dp];
#pragma unroll
        for (int q = 0; q < Dp; ++q) xq[q] = __shfl(xv[q], src);
        float best = -INFINITY;
        int bj = 0x7fffffff;
        for (int j = lane; j < k; j += 64) {
          float sc = -0.5f * ncs[j];
          const float *cr = Cf + (size_t)j * cs;
#pragma unroll
          for (int q = 0; q < Dp; ++q)
            if (q < d) sc = __builtin_fmaf(xq[q], cr[q], sc);
          if (sc > best) { best = sc; bj = j; }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          const float ov = __shfl_xor(best, o);
          const int oj = __shfl_xor(bj, o);
          if (ov > best || (ov == best && oj < bj)) { best = ov; bj = oj; }
        }
        if (lane == 0) set_label(labels, base + pp, bj, imin + (size_t)inst * k, (int)(w.y + pp));
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0 && exact_ct) atomicAdd(nexact, exact_ct);
  if (tid == 0 && nkp && kp_ct) atomicAdd(nkp + 2, kp_ct);
}

// Left-to-right sum of m FP64 values at stride st, in chunks of KM_CHUNK
// whose sums are again added left to right (m <= KM_CHUNK^2).
__device__ inline double csum2(const double *v, int m, int st) {
  if (m <= KM_CHUNK) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += v[(long)i * st];
    return s;
  }
  double tot = 0.0;
  for (int a = 0; a < m; a += KM_CHUNK) {
    double s = 0.0;
    const int e = min(m, a + KM_CHUNK);
    for (int i = a; i < e; ++i) s += v[(long)i * st];
    tot += s;
  }
  return tot;
}

// Exclusive scan of v[0..n) in LDS by the whole block (a multiple of 64
// threads, at most 1024); tmp holds at least 65 ints.  Each thread scans a
// run of ceil(n / threads) entries; the runs' totals are scanned by wave
// shuffles, then the waves' totals.  Returns the total (integer sums: any
// order gives the same values).
__device__ int block_scan_any(int *v, int n, int *tmp) {
  const int t = threadIdx.x, nt = blockDim.x, lane = t & 63, w = t >> 6, nw = nt >> 6;
  const int per = (n + nt - 1) / nt, i0 = min(n, t * per), i1 = min(n, i0 + per);
  int s = 0;
  for (int i = i0; i < i1; ++i) s += v[i];
  int x = s;  // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  if (t < 64) {
    const int wt = t < nw ? tmp[t] : 0;
    int z = wt;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(z, o);
      if (t >= o) z += y;
    }
    if (t < nw) tmp[16 + t] = z - wt;
    if (t == nw - 1) tmp[64] = z;
  }
  __syncthreads();
  int acc = tmp[16 + w] + x - s;
  for (int i = i0; i < i1; ++i) {
    const int y = v[i];
    v[i] = acc;
    acc += y;
  }
  __syncthreads();
  const int total = tmp[64];
  __syncthreads();
  return total;
}

// Update 1/3, persistent over instances (largest first), one 16-wave
// workgroup per instance: stable counting sort of the instance's points by
// label.  A first pass counts the labels (per wave and label, then per
// label).  Staged (STEPS > 0, k up to ~480): the points are then placed in
// tiles of 1024 * STEPS consecutive rows: each wave ranks its 64 * STEPS rows
// of the tile among equal labels (LDS lane masks per label class, label mod
// 64, plus ballots on the remaining bits), the per-(wave, label) counts are
// scanned over waves and
// labels, every row goes to its label-major slot of the tile in LDS, and the
// block writes the tile out in slot order: a label's members of one tile are
// consecutive in perm, so the writes come in runs of ~1024 * STEPS / k entries
// instead of one scattered 4-byte write per point.  Otherwise wave w owns a
// contiguous segment of the points; per (wave, label) counts give every wave
// its own cursors, and each wave scatters its segment in order with no block
// barrier (ranks among equal labels from ballots over the label bits).
// Writes perm (instance-local indices, label-major), counts, offsets and the
// chunk prefix of every cluster.
// With reuse (rstate = [the other parity holds the previous update of these
// points, its parity]): each cluster's first member position at which this
// sort differs from the previous one goes to fd, for k_batch_chunks_wave.
// Labels are read ahead of their use (eight per lane in the count pass, the
// next tile's STEPS in the staged pass).  The staged pass issues exactly
// STEPS stores per thread and tile (rows past the end write a dummy
// area past the perm arrays, dummy[0..1024)), so the loads of the next tile,
// issued before them, are waited for without waiting for the stores.
// Parity of an update: the one the previous update did not write (rstate[1]
// once rstate[0] says it wrote one); the arrays come in pairs.
__device__ inline int upd_parity(const int *rstate) { return rstate[0] ? rstate[1] ^ 1 : 0; }

template <int STEPS>  // staged pass: 64-row steps per wave and tile (4 or 8); 0: unstaged
__global__ __launch_bounds__(1024) void k_batch_sort(
    int nb, const int *__restrict__ order, const long *__restrict__ xoff,
    const int *__restrict__ nvec, int k, const int *__restrict__ labels, int *perm0, int *perm1,
    long tot, int *gcnt0, int *gcnt1, int *goff0, int *goff1, int *gcpre0, int *gcpre1,
    const long *__restrict__ choff, int4 *__restrict__ desc, int *__restrict__ counter,
    const int *__restrict__ rstate, int *__restrict__ fd, int *__restrict__ imin) {
  // cur[16][k], offs[k], cpre[k], cnt[k], fdc[k], ims[k]; STAGED: then the
  // tile's slots stage[TILE] (local row | label << SH) and per wave the
  // masks of the lanes holding each label class in the current step (all
  // zero between steps)
  constexpr bool STAGED = STEPS > 0;
  constexpr int NS = STAGED ? STEPS : 1;  // register arrays of the staged pass
  constexpr int TILE = 16 * 64 * (STAGED ? STEPS : 1);  // rows per tile
  constexpr int SH = STEPS == 8 ? 13 : 12;              // bits of a row within a tile
  extern __shared__ __attribute__((aligned(8))) int sh[];
  int *cur = sh, *offs = sh + 16 * k, *cpre = offs + k, *cnts = cpre + k, *fdc = cnts + k,
      *ims = fdc + k;
  int *stage = sh + ((21 * k + 1) & ~1);
  unsigned long long *pm = reinterpret_cast<unsigned long long *>(stage + TILE);
  const bool reuse = rstate[0] != 0;
  const int par = upd_parity(rstate);
  int *__restrict__ perm = par ? perm1 : perm0;
  int *__restrict__ dummy = perm + tot;  // 1024 ints past the points
  int *__restrict__ gcnt = par ? gcnt1 : gcnt0;
  int *__restrict__ goff = par ? goff1 : goff0;
  int *__restrict__ gcpre = par ? gcpre1 : gcpre0;
  __shared__ int tmp[65];
  __shared__ int item_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int nbits = k > 1 ? 32 - __clz(k - 1) : 0;
  for (;;) {
    if (t == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= nb) break;
    const int inst = order[it];
    const int n = nvec[inst];  // instance sizes stay below 2^31
    const long base = xoff[inst];
    const int *lab = labels + base;
    int *pb = perm + base;
    for (int e = t; e < 16 * k; e += blockDim.x) cur[e] = 0;
    for (int j = t; j < k; j += blockDim.x) {
      fdc[j] = 0;
      ims[j] = reuse ? imin[(size_t)inst * k + j] : 0;
    }
    if (STAGED)
      for (int e = t; e < 16 * 64; e += blockDim.x) pm[e] = 0ull;
    __syncthreads();
    const int S = ((n + 15) / 16 + 63) / 64 * 64;
    const int a = min(n, w * S), e = min(n, a + S);
    // per-(wave, label) counts; fdc[j] = members of j below its smallest
    // changed row, i.e. the length of j's unchanged member prefix (the list
    // is in ascending row order).  Loads are clamped into the segment, not
    // branched around.
    for (int p0 = a; p0 < e; p0 += 512) {
      int Lv[8], iv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) Lv[u] = lab[min(p0 + u * 64 + lane, e - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) iv[u] = ims[Lv[u]];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int p = p0 + u * 64 + lane;
        if (p < e) {
          atomicAdd(&cur[w * k + Lv[u]], 1);
          if (p < iv[u]) atomicAdd(&fdc[Lv[u]], 1);
        }
      }
    }
    __syncthreads();
    for (int j = t; j < k; j += blockDim.x) {
      int c = 0;
      for (int ww = 0; ww < 16; ++ww) c += cur[ww * k + j];
      offs[j] = c;
      cnts[j] = c;
      cpre[j] = (c + KM_CHUNK - 1) / KM_CHUNK;
      gcnt[(size_t)inst * k + j] = c;
    }
    __syncthreads();
    block_scan_any(offs, k, tmp);
    const int nch = block_scan_any(cpre, k, tmp);
    for (int j = t; j < k; j += blockDim.x) {
      int c = offs[j];
      for (int ww = 0; ww < 16; ++ww) {  // STAGED: counts of the first tile
        const int h = cur[ww * k + j];
        cur[ww * k + j] = STAGED ? 0 : c;
        c += h;
      }
      goff[(size_t)inst * k + j] = offs[j];
      gcpre[(size_t)inst * (k + 1) + j] = cpre[j];
      if (reuse) fd[(size_t)inst * k + j] = fdc[j];
      imin[(size_t)inst * k + j] = 0x7fffffff;  // cleared for the next iteration
    }
    if (t == 0) gcpre[(size_t)inst * (k + 1) + k] = nch;
    __syncthreads();
    // chunk descriptors (instance, first and end member position, chunk and
    // cluster); the instance's unused slots get empty ranges
    {
      int4 *dd = desc + choff[inst];
      const int cap = (int)(choff[inst + 1] - choff[inst]);
      for (int j = t; j < k; j += blockDim.x) {
        const int o = offs[j], nj = cnts[j];
        int q = 0;
        for (int c = cpre[j], a0 = o; a0 < o + nj; ++c, a0 += KM_CHUNK, ++q)
          dd[c] = make_int4(inst, a0, min(o + nj, a0 + KM_CHUNK), q * 1024 + j);
      }
      for (int c = nch + t; c < cap; c += blockDim.x) dd[c] = make_int4(0, 0, 0, 0);
    }
    if (STAGED) {
      // the label cursors (instance-local perm positions) of the current and
      // the next tile alternate between offs and cpre; cnts holds the tile's
      // label offsets (the descriptors above were the last use of both)
      int *toff = cnts;
      int *cw = cur + w * k;
      unsigned long long *pw = pm + w * 64;
      const int wr = w * (64 * STEPS);  // the wave's first row in a tile
      int Lc[NS], Ln[NS];
      auto fetch = [&](int t0, int (&v)[NS]) {  // clamped into the instance
#pragma unroll
        for (int u = 0; u < STEPS; ++u) v[u] = lab[min(t0 + wr + u * 64 + lane, n - 1)];
      };
      // labels are clamped into [0, k) where they arrive in registers: this
      // also waits for the first tile's loads here, outside the tile loop
      if (n > 0) {
        fetch(0, Lc);
#pragma unroll
        for (int u = 0; u < STEPS; ++u) Lc[u] = min(max(Lc[u], 0), k - 1);
      }
      int tp = 0;
      for (int t0 = 0; t0 < n; t0 += TILE, tp ^= 1) {
        fetch(t0 + TILE, Ln);  // unconditional: row n - 1 past the end
        // ranks of the wave's rows among its rows of equal label (row order)
        int lr[NS];
#pragma unroll
        for (int u = 0; u < STEPS; ++u) {
          const int p = t0 + wr + u * 64 + lane;
          const bool ok = p < n;
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
          const int c = ok ? cw[L] : 0;
          lr[u] = c + __popcll(m & below);
          __builtin_amdgcn_wave_barrier();
          if (ok && (m >> lane) == 1ull) cw[L] = c + __popcll(m);  // the label's last lane
          if (ok && (cls >> lane) == 1ull) pw[L & 63] = 0ull;       // the class's last lane
          __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // per label: exclusive prefix of the counts over waves (in place),
        // the next tile's cursor, and the exclusive prefix of the tile's
        // label totals over labels (k <= 1024: one label per thread)
        int s = 0;
        if (t < k) {
          for (int ww = 0; ww < 16; ++ww) {
            const int c = cur[ww * k + t];
            cur[ww * k + t] = s;
            s += c;
          }
          (tp ? offs : cpre)[t] = (tp ? cpre : offs)[t] + s;
        }
        int x = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(x, o);
          if (lane >= o) x += y;
        }
        if (lane == 63) tmp[w] = x;
        __syncthreads();
        if (t < 64) {
          const int wt = t < 16 ? tmp[t] : 0;
          int z = wt;
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(z, o);
            if (t >= o) z += y;
          }
          if (t < 16) tmp[16 + t] = z - wt;
        }
        __syncthreads();
        if (t < k) toff[t] = tmp[16 + w] + x - s;
        __syncthreads();
        // every row to its label-major slot of the tile
#pragma unroll
        for (int u = 0; u < STEPS; ++u) {
          const int L = Lc[u];
          const int r = wr + u * 64 + lane;
          if (t0 + r < n) stage[toff[L] + cw[L] + lr[u]] = r | (L << SH);
        }
        __syncthreads();
        // the wave's counters back to zero for the next tile (only this
        // wave writes them before the next tile's first barrier)
        for (int j = lane; j < k; j += 64) cw[j] = 0;
        // the tile in slot order: runs of one label go to consecutive
        // positions from the label's cursor
        const int rows = min(TILE, n - t0);
        const int *gcc = tp ? cpre : offs;
#pragma unroll
        for (int q = 0; q < TILE / 1024; ++q) {
          const int i = t + q * 1024;
          const bool ok = i < rows;
          const int v = stage[ok ? i : 0];
          const int L = v >> SH;
          int *dst = ok ? pb + gcc[L] + (i - toff[L]) : dummy + t;
          *dst = t0 + (v & (TILE - 1));
        }
#pragma unroll
        for (int u = 0; u < STEPS; ++u) Lc[u] = min(max(Ln[u], 0), k - 1);
      }
      __syncthreads();
      continue;
    }
    int *cw = cur + w * k;
    for (int p0 = a; p0 < e; p0 += 512) {
      int Lv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) Lv[u] = lab[min(p0 + u * 64 + lane, e - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (p0 + u * 64 >= e) break;  // uniform over the wave
        const int p = p0 + u * 64 + lane;
        const bool valid = p < e;
        const int L = Lv[u];
        // the lanes of this step holding label L: a match over the label bits
        uint64_t m = __ballot(valid);
        for (int bt = 0; bt < nbits; ++bt) {
          const uint64_t bb = __ballot(valid && ((L >> bt) & 1));
          m &= ((L >> bt) & 1) ? bb : ~bb;
        }
        if (!valid) m = 0;
        const int cwL = cw[L];
        const int rank = __popcll(m & below);
        const int pos = cwL + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid) pb[pos] = p;
        if (valid && rank == 0) cw[L] = pos + __popcll(m);
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
  }
}

// Update 2/3: one wave per chunk slot (d <= 64).  The 64 lanes gather 64
// members' rows into LDS at once; then lane q < d adds dimension q and its
// square (exact in FP64), member by member in ascending order, so the sums
// are the same sequential FP64 sums.  The squared-norm column is the
// per-dimension square sums added in ascending dimension order.
// With reuse (as k_batch_sort), a chunk whose members all lie before its
// cluster's first changed position copies the previous update's partials.
template <int DT>
__global__ __launch_bounds__(256) void k_batch_chunks_wave(
    long slots, const int4 *__restrict__ desc, const long *__restrict__ xoff, int d,
    const float *__restrict__ X, const int *perm0, const int *perm1, double *part0, double *part1,
    int nb, int k,
    const long *__restrict__ choff, const int *__restrict__ fd, const int *gcnt0,
    const int *gcnt1, const int *gcpre0, const int *gcpre1, const int *__restrict__ rstate) {
  __shared__ float xs[4][64][DT + 1];
  __shared__ double ts[4][DT];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long s = (long)blockIdx.x * 4 + w;
  if (s >= slots) return;
  const int par = upd_parity(rstate);
  const int *__restrict__ perm = par ? perm1 : perm0;
  double *__restrict__ part = par ? part1 : part0;
  const double *__restrict__ part_prev = par ? part0 : part1;
  const int *__restrict__ gcnt_prev = par ? gcnt0 : gcnt1;
  const int *__restrict__ gcpre_prev = par ? gcpre0 : gcpre1;
  const int4 dsc = desc[s];
  const int a = dsc.y, e = dsc.z;
  if (a >= e) return;
  const int lo = dsc.x;  // the slot's instance
  const long base = xoff[lo];
  if (rstate[0]) {
    const int j = dsc.w & 1023, q = dsc.w >> 10;
    const int q0 = q * KM_CHUNK, end_new = q0 + (e - a);
    const int end_old = min(q0 + KM_CHUNK, gcnt_prev[(size_t)lo * k + j]);
    if (end_new == end_old && end_new <= fd[(size_t)lo * k + j]) {
      const long src = choff[lo] + gcpre_prev[(size_t)lo * (k + 1) + j] + q;
      if (lane <= d) part[s * (d + 1) + lane] = part_prev[src * (d + 1) + lane];
      return;
    }
  }
  double acc = 0.0, sq = 0.0;
  for (int r0 = a; r0 < e; r0 += 64) {
    const int r = r0 + lane;
    if (r < e) {
      const long idx = base + perm[base + r];
      if (d == DT && (DT & 3) == 0) {
        const float4 *x4 = reinterpret_cast<const float4 *>(X + idx * DT);
#pragma unroll
        for (int v = 0; v < DT / 4; ++v) {
          const float4 xv = x4[v];
          xs[w][lane][4 * v] = xv.x;
          xs[w][lane][4 * v + 1] = xv.y;
          xs[w][lane][4 * v + 2] = xv.z;
          xs[w][lane][4 * v + 3] = xv.w;
        }
      } else {
        const float *x = X + idx * d;
#pragma unroll
        for (int qq = 0; qq < DT; ++qq)
          if (qq < d) xs[w][lane][qq] = x[qq];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int m = min(64, e - r0);
    if (lane < d) {
      for (int i = 0; i < m; ++i) {
        const double v = xs[w][i][lane];
        acc += v;
        sq = __builtin_fma(v, v, sq);  // the square of a float is exact in FP64
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane < d) {
    part[s * (d + 1) + lane] = acc;
    ts[w][lane] = sq;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane == 0) {
    double qv = 0.0;
    for (int u = 0; u < d; ++u) qv += ts[w][u];
    part[s * (d + 1) + d] = qv;
  }
}

// Update 3/3: one workgroup per instance: canonical combination of the chunk
// sums, new centres (empty clusters keep theirs), ||c||^2 chains,
// compactness from the moments against the new centres.
// FP32 chain over the dimensions of (c - bf16(c))^2 (each residual exact)
__device__ inline float bf16_residual_sq(const float *c, int d) {
  float t = 0.0f;
  for (int q = 0; q < d; ++q) {
    const float v = c[q];
    const float r = v - __uint_as_float((uint32_t)b_bf16(v) << 16);
    t = __builtin_fmaf(r, r, t);
  }
  return t;
}

// block-wide maxima of two values (256 threads), written by thread 0
__device__ inline void block_max2(float a, float b, float *ra, float *rb, float2 *out) {
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      ra[threadIdx.x] = fmaxf(ra[threadIdx.x], ra[threadIdx.x + o]);
      rb[threadIdx.x] = fmaxf(rb[threadIdx.x], rb[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_float2(ra[0], rb[0]);
}

// Also each centre's drift (an upper bound of how far it moved, 0 exactly for
// an unmoved centre) and the instance's largest drift.
// The clusters go in slabs whose sums fit in LDS: first every (cluster,
// column) sum of the slab and the old centre values, one per thread, so the
// loads of the whole slab are in flight together; then one thread per
// cluster finishes it from LDS (the same sums and chains as before, in the
// same orders).
#define B_FIN_LDS 4352  // FP64 sums per slab (256 clusters at d = 16)
__global__ __launch_bounds__(256) void k_batch_finalize(
    int d, int k, const long *__restrict__ choff, const int *gcnt0, const int *gcnt1,
    const int *gcpre0, const int *gcpre1, const double *part0, const double *part1,
    float *__restrict__ C, float *__restrict__ nc, float2 *__restrict__ bnd,
    double *__restrict__ comp_out, int *__restrict__ empty_out, float *__restrict__ drift,
    float *__restrict__ dmaxv, const int *__restrict__ rstate) {
  __shared__ double Ss[B_FIN_LDS];
  __shared__ float Co[B_FIN_LDS];
  __shared__ double cj[B_MAXK];
  __shared__ int nempty;
  __shared__ float rx[256], ry[256];
  float mx = 0.0f, dm = 0.0f, md = 0.0f;
  const int par = upd_parity(rstate);
  const int *__restrict__ gcnt = par ? gcnt1 : gcnt0;
  const int *__restrict__ gcpre = par ? gcpre1 : gcpre0;
  const double *__restrict__ part = par ? part1 : part0;
  const int inst = blockIdx.x, t = threadIdx.x, ld = d + 1;
  if (t == 0) nempty = 0;
  const double *pi = part + choff[inst] * ld;
  const int *cp = gcpre + (size_t)inst * (k + 1);
  float *Ci = C + (size_t)inst * k * d;
  const int JS = min(k, B_FIN_LDS / ld);
  for (int j0 = 0; j0 < k; j0 += JS) {
    const int jn = min(JS, k - j0);
    for (int e = t; e < jn * ld; e += blockDim.x) {
      const int jj = e / ld, qq = e - jj * ld, j = j0 + jj;
      const int c0 = cp[j];
      Ss[e] = csum2(pi + (long)c0 * ld + qq, cp[j + 1] - c0, ld);
      if (qq < d) Co[jj * d + qq] = Ci[(size_t)j * d + qq];
    }
    __syncthreads();
    for (int j = j0 + t; j < j0 + jn; j += blockDim.x) {
      const double *Sj = Ss + (j - j0) * ld;
      const float *Cold = Co + (j - j0) * d;
      const int nj = gcnt[(size_t)inst * k + j];
      double a = 0.0, b = 0.0;
      float s2 = 0.0f, dq = 0.0f, dmx = 0.0f, res = 0.0f;
      for (int qq = 0; qq < d; ++qq) {
        const double S = Sj[qq];
        const float old = Cold[qq];
        float cv = old;
        if (nj > 0) cv = (float)(S / (double)nj);
        Ci[(size_t)j * d + qq] = cv;
        const float tq = cv - old;
        dq = __builtin_fmaf(tq, tq, dq);
        const float at = fabsf(tq);
        if (!(at <= dmx)) dmx = at;  // NaN propagates
        s2 = __builtin_fmaf(cv, cv, s2);
        // FP32 chain of (c - bf16(c))^2 over the dimensions (each residual exact)
        const float r = cv - __uint_as_float((uint32_t)b_bf16(cv) << 16);
        res = __builtin_fmaf(r, r, res);
        const double cd = (double)cv;
        a += cd * S;
        b += cd * cd;
      }
      const double Q = Sj[d];
      nc[(size_t)inst * k + j] = s2;
      mx = fmaxf(mx, s2);
      dm = fmaxf(dm, res);
      cj[j] = (Q - 2.0 * a) + (double)nj * b;
      if (nj == 0) atomicAdd(&nempty, 1);
      const float dv = drift_bound(dq, dmx, d);
      drift[(size_t)inst * k + j] = dv;
      if (!(dv <= md)) md = dv;
    }
    __syncthreads();
  }
  if (t == 0) {
    comp_out[inst] = csum2(cj, k, 1);
    empty_out[inst] = nempty;
  }
  block_max2(mx, dm, rx, ry, bnd + inst);
  __syncthreads();
  rx[t] = md;  // the instance's largest drift (NaN propagates)
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o && !(rx[t + o] <= rx[t])) rx[t] = rx[t + o];
    __syncthreads();
  }
  if (t == 0) dmaxv[inst] = rx[0];
}

// After an update: the drift bounds hold (for the BF16 assign) and the
// parity it wrote becomes the previous one (a launch of its own, since every
// block of the update's kernels reads the parity).
__global__ void k_batch_flip(int *__restrict__ rstate, int *__restrict__ bvalid) {
  rstate[1] = upd_parity(rstate);
  rstate[0] = 1;
  *bvalid = 1;
}

// Per instance: an upper bound of the largest |x| of its points (one block
// per instance), for the drift bounds' FP32 chain error.
__global__ __launch_bounds__(256) void k_batch_xmax(int d, const long *__restrict__ xoff,
                                                    const int *__restrict__ nvec,
                                                    const float *__restrict__ X,
                                                    float *__restrict__ xmax) {
  __shared__ float rx[256];
  const int inst = blockIdx.x, t = threadIdx.x;
  const float *Xi = X + xoff[inst] * d;
  float m = 0.0f;
  for (long p = t; p < nvec[inst]; p += 256) {
    float sq = 0.0f;
    for (int q = 0; q < d; ++q) sq = __builtin_fmaf(Xi[p * d + q], Xi[p * d + q], sq);
    const float xn = sqrtf(sq * (1.0f + (float)d * 1.2e-7f + 1e-6f)) * (1.0f + 1e-6f) + 1e-30f;
    if (!(xn <= m)) m = xn;
  }
  rx[t] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o && !(rx[t + o] <= rx[t])) rx[t] = rx[t + o];
    __syncthreads();
  }
  if (t == 0) xmax[inst] = rx[0];
}

// Per instance: the largest ||c_j||^2 (nc, already computed) and the largest
// ||c_j - bf16(c_j)||^2, for centres set from outside the iteration.
__global__ __launch_bounds__(256) void k_batch_bounds(int d, int k, const float *__restrict__ C,
                                                      const float *__restrict__ nc,
                                                      float2 *__restrict__ bnd) {
  __shared__ float rx[256], ry[256];
  const int inst = blockIdx.x;
  float mx = 0.0f, dm = 0.0f;
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    mx = fmaxf(mx, nc[(size_t)inst * k + j]);
    dm = fmaxf(dm, bf16_residual_sq(C + ((size_t)inst * k + j) * d, d));
  }
  block_max2(mx, dm, rx, ry, bnd + inst);
}

// ||c||^2 chains of every instance's centres.
__global__ void k_batch_norms(long kk, int d, const float *__restrict__ C, float *__restrict__ nc) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kk) return;
  float s = 0.0f;
  for (int q = 0; q < d; ++q) s = __builtin_fmaf(C[j * d + q], C[j * d + q], s);
  nc[j] = s;
}

// First k points of each instance as its centres.
__global__ void k_batch_first_k(int nb, int k, int d, const long *__restrict__ xoff,
                                const float *__restrict__ X, float *__restrict__ C) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nb * k * d) return;
  const long inst = e / ((long)k * d), r = e - inst * k * d;
  C[e] = X[xoff[inst] * d + r];
}

// ------------------------------------------------------------------ host
#define BCHK(x, err)                                                              \
  do {                                                                            \
    hipError_t _e = (x);                                                          \
    if (_e != hipSuccess) {                                                       \
      fprintf(stderr, "kmeans_amd: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__, \
              hipGetErrorString(_e));                                             \
      return err;                                                                 \
    }                                                                             \
  } while (0)

static int pick_dt(int d) {
  const int dts[] = {4, 8, 16, 32, 64};
  for (int v : dts)
    if (d <= v) return v;
  return 0;
}

template <int DT>
static hipError_t launch_assign_dt(Batch *b, int P, int nitems, const int2 *items) {
  const size_t lds = sizeof(float) * ((size_t)b->k * DT + b->k);
  static size_t attr = 64 * 1024;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_batch_assign_f32<DT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  const int grid = num_cus() * 4;
  hipLaunchKernelGGL(k_batch_assign_f32<DT>, dim3(grid), dim3(256), lds, b->stream, nitems, items,
                     P, b->xoff, b->nvec, b->d, b->k, b->X, b->C, b->nc, b->labels, b->counters,
                     b->imin);
  return hipGetLastError();
}

template <int DT>
static hipError_t launch_assign_x(Batch *b, int P, int nitems, const int2 *items) {
  const size_t lds = sizeof(float) * ((size_t)b->k * DT + b->k);
  static size_t attr = 64 * 1024;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_batch_assign_f32x<DT, 4>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  const int grid = num_cus() * 4;
  hipLaunchKernelGGL((k_batch_assign_f32x<DT, 4>), dim3(grid), dim3(256), lds, b->stream, nitems,
                     items, P, b->xoff, b->nvec, b->k, b->X, b->C, b->nc, b->labels, b->counters,
                     b->imin);
  return hipGetLastError();
}

static size_t bf16_lds(const Batch *b) {
  const int Kp = (b->k + 31) / 32 * 32, Dp = b->d <= 16 ? 16 : 32;
  return (size_t)Kp * Dp * 2 + (size_t)Kp * 4 + (size_t)b->k * (b->d | 1) * 4 + (size_t)b->k * 4;
}

static hipError_t launch_assign(Batch *b, int P, int nitems, const int2 *items) {
  if (b->opt.precision == KM_PREC_BF16) {
    const size_t lds = bf16_lds(b);
    const int grid = num_cus() * 4;
    if (lds > 64 * 1024) return hipErrorInvalidValue;  // batch_create keeps it within 64 KB
    if (b->xb)
      hipLaunchKernelGGL(k_batch_classify, dim3(nitems), dim3(256),
                         sizeof(float) * (size_t)b->k * ((b->d | 1) + 2), b->stream, nitems, items, P,
                         b->xoff, b->nvec, b->d, b->k, b->X, b->C, b->nc, b->bnd, b->labels, b->xb,
                         b->bvalid, b->drift, b->dmaxv, b->xmax, b->nkp, b->imin, b->olist,
                         b->ocnt);
#define BA(KS)                                                                                   \
  hipLaunchKernelGGL(k_batch_assign_bf16<KS>, dim3(grid), dim3(256), lds, b->stream, nitems, items, \
                     P, b->xoff, b->nvec, b->d, b->k, b->X, b->C, b->nc, b->bnd, b->labels,        \
                     b->counters, b->nexact, FLT_MAX, b->xb, b->bvalid, b->drift, b->dmaxv,        \
                     b->xmax, b->nkp, b->imin, b->olist, b->ocnt)
    if (b->d <= 16) BA(1);
    else BA(2);
#undef BA
    return hipGetLastError();
  }
  if (b->d == 16) return launch_assign_x<16>(b, P, nitems, items);
  if (b->d == 8) return launch_assign_x<8>(b, P, nitems, items);
  if (b->d == 32) return launch_assign_x<32>(b, P, nitems, items);
  switch (pick_dt(b->d)) {
    case 4: return launch_assign_dt<4>(b, P, nitems, items);
    case 8: return launch_assign_dt<8>(b, P, nitems, items);
    case 16: return launch_assign_dt<16>(b, P, nitems, items);
    case 32: return launch_assign_dt<32>(b, P, nitems, items);
    default: return launch_assign_dt<64>(b, P, nitems, items);
  }
}

static hipError_t launch_update(Batch *b, int slot) {
  double *co = b->comp + (size_t)slot * b->nb;
  int *eo = b->empty + (size_t)slot * b->nb;
  // the staged pass adds the tile's slots and 16 x 64 class masks; 8192-row
  // tiles while everything stays within 64 KB (k up to ~290), else 4096-row
  // tiles (k up to ~480), else the unstaged pass
  const size_t lds0 = sizeof(int) * (size_t)((21 * b->k + 1) & ~1);
  const size_t masks = sizeof(unsigned long long) * 16 * 64;
  const size_t lds8 = lds0 + sizeof(int) * 8192 + masks, lds4 = lds0 + sizeof(int) * 4096 + masks;
  const int steps = lds8 <= 64 * 1024 ? 8 : lds4 <= 64 * 1024 ? 4 : 0;
  const size_t lds = steps == 8 ? lds8 : steps == 4 ? lds4 : lds0;
  static size_t attr = 64 * 1024;
  if (steps == 0 && lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_batch_sort<0>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL(steps == 8 ? k_batch_sort<8> : steps == 4 ? k_batch_sort<4> : k_batch_sort<0>,
                     dim3(std::min(b->nb, num_cus() * 2)), dim3(1024), lds, b->stream, b->nb,
                     b->order, b->xoff, b->nvec, b->k, b->labels, b->perm[0], b->perm[1],
                     b->total, b->gcnt[0], b->gcnt[1], b->goff[0], b->goff[1], b->gcpre[0], b->gcpre[1],
                     b->choff, b->desc, b->counters + 1, b->rstate, b->fd, b->imin);
#define CHW(DT)                                                                                  \
  hipLaunchKernelGGL(k_batch_chunks_wave<DT>, dim3((unsigned)((b->slots + 3) / 4)), dim3(256), 0,  \
                     b->stream, b->slots, b->desc, b->xoff, b->d, b->X, b->perm[0], b->perm[1],      \
                     b->part[0],                                                                   \
                     b->part[1], b->nb, b->k, b->choff, b->fd, b->gcnt[0], b->gcnt[1], b->gcpre[0], \
                     b->gcpre[1], b->rstate)
  switch (pick_dt(b->d)) {
    case 4: CHW(4); break;
    case 8: CHW(8); break;
    case 16: CHW(16); break;
    case 32: CHW(32); break;
    default: CHW(64); break;
  }
#undef CHW
  hipLaunchKernelGGL(k_batch_finalize, dim3(b->nb), dim3(256), 0, b->stream, b->d, b->k, b->choff,
                     b->gcnt[0], b->gcnt[1], b->gcpre[0], b->gcpre[1], b->part[0], b->part[1], b->C,
                     b->nc, b->bnd, co, eo, b->drift, b->dmaxv, b->rstate);
  hipLaunchKernelGGL(k_batch_flip, dim3(1), dim3(1), 0, b->stream, b->rstate, b->bvalid);
  return hipGetLastError();
}

static void build_items(const Batch *b, int P, std::vector<int2> &items) {
  items.clear();
  // largest instances first so the tail of the queue is short items
  std::vector<int> ord(b->nb);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return b->n_h[a] > b->n_h[c]; });
  for (int i : ord)
    for (long s = 0; s < b->n_h[i]; s += P) items.push_back(make_int2(i, (int)s));
}

static hipError_t upload_items(Batch *b, int P) {
  std::vector<int2> items;
  build_items(b, P, items);
  hipFree(b->items);
  b->items = nullptr;
  hipError_t e = hipMalloc(&b->items, sizeof(int2) * std::max<size_t>(items.size(), 1));
  if (e != hipSuccess) return e;
  e = hipMemcpy(b->items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice);
  b->nitems = (int)items.size();
  b->P = P;
  if (e == hipSuccess && b->xb && items.size() > b->ocnt_cap) {  // one open count per item
    hipFree(b->ocnt);
    b->ocnt = nullptr;
    b->ocnt_cap = 0;
    if ((e = hipMalloc(&b->ocnt, sizeof(int) * items.size())) == hipSuccess)
      b->ocnt_cap = items.size();
  }
  return e;
}

// ---- autotuner: points per work item of the assign phase, timed on the
// caller's data and cached per shape (in process and in a text file named by
// KMEANS_AMD_TUNE_CACHE, default $HOME/.cache/kmeans_amd_tune.txt).
static std::mutex g_tune_mu;
static std::map<std::string, int> g_tune;
static bool g_tune_loaded = false;

static std::string tune_path() {
  const char *p = getenv("KMEANS_AMD_TUNE_CACHE");
  if (p && *p) return p;
  const char *h = getenv("HOME");
  return std::string(h && *h ? h : "/tmp") + "/.cache/kmeans_amd_tune.txt";
}

static std::string tune_key(const Batch *b) {
  int lg_n = 0, lg_b = 0;
  while ((1L << lg_n) < b->total) ++lg_n;
  while ((1 << lg_b) < b->nb) ++lg_b;
  std::ostringstream s;
  s << "batch_assign " << (b->opt.precision == KM_PREC_BF16 ? "bf16" : "f32") << " v2 d=" << b->d << " k=" << b->k << " lgN=" << lg_n << " lgB=" << lg_b;
  return s.str();
}

static int autotune(Batch *b) {
  const std::string key = tune_key(b);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    if (!g_tune_loaded) {
      std::ifstream f(tune_path());
      std::string line;
      while (std::getline(f, line)) {
        const size_t eq = line.rfind('=');
        const size_t sp = line.rfind(' ');
        if (sp == std::string::npos || eq == std::string::npos || eq < sp) continue;
        g_tune[line.substr(0, sp)] = atoi(line.c_str() + eq + 1);
      }
      g_tune_loaded = true;
    }
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      b->stats.tuned_from_cache = 1;
      return it->second;
    }
  }
  const int cands[] = {256, 512, 1024, 2048, 4096};
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int best = 1024;
  float best_ms = 1e30f;
  for (int ci = 0; ci < 5; ++ci) {
    const int P = cands[ci];
    if (upload_items(b, P) != hipSuccess) continue;
    float tot = 0.0f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(64), 0, b->stream, b->counters, 2, 0);
      hipEventRecord(e0, b->stream);
      launch_assign(b, P, b->nitems, b->items);
      hipEventRecord(e1, b->stream);
      hipEventSynchronize(e1);
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      if (rep) tot += ms;  // first run warms up
    }
    b->stats.tune_ms[ci] = tot / 2.0f;
    if (tot < best_ms) { best_ms = tot; best = P; }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = best;
    const std::string path = tune_path();
    const size_t sl = path.rfind('/');
    if (sl != std::string::npos) {
      std::string dir = path.substr(0, sl);
      std::string cmd = "mkdir -p '" + dir + "' 2>/dev/null";
      (void)system(cmd.c_str());
    }
    std::ofstream f(path, std::ios::app);
    if (f) f << key << " P=" << best << "\n";
  }
  return best;
}

static void batch_free_graph(Batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_iters = 0;
}

static void batch_free(Batch *b) {
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_free_graph(b);
  for (hipEvent_t ev : b->gev) hipEventDestroy(ev);
  b->gev.clear();
  void *bufs[] = {b->X_owned, b->xoff, b->nvec, b->C, b->nc, b->labels,
                  b->comp, b->choff, b->empty, b->order, b->items, b->counters,
                  b->desc, b->nexact, b->bnd,
                  b->xb, b->drift, b->dmaxv, b->xmax, b->bvalid, b->nkp, b->fd, b->rstate,
                  b->imin, b->olist, b->ocnt};
  for (void *p : bufs) hipFree(p);
  for (int q = 0; q < 2; ++q) {
    hipFree(b->perm[q]);
    hipFree(b->part[q]);
    hipFree(b->gcnt[q]);
    hipFree(b->goff[q]);
    hipFree(b->gcpre[q]);
  }
  if (b->own_stream && b->stream) hipStreamDestroy(b->stream);
}

static kmeansErrorStatus batch_create(Batch *b, int nb, const long *n, int d, int k,
                                      const kmeansOptions *o) {
  if (nb <= 0 || !n || d <= 0 || d > B_MAXD || k <= 0 || k > B_MAXK)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  if (o) b->opt = *o;
  if ((b->opt.precision != KM_PREC_FP32 && b->opt.precision != KM_PREC_BF16) ||
      b->opt.proposer != KM_PROPOSER_NONE || b->opt.tile_rows)
    return KMEANS_ERROR_INVALID_ARGUMENT;
  b->d = d;
  b->k = k;
  if (b->opt.precision == KM_PREC_BF16 && (d > 32 || bf16_lds(b) > 64 * 1024))
    return KMEANS_ERROR_INVALID_ARGUMENT;
  const int dt = pick_dt(d);
  if ((size_t)k * (dt + 1) * sizeof(float) > 64 * 1024) return KMEANS_ERROR_INVALID_ARGUMENT;
  b->nb = nb;
  b->d = d;
  b->k = k;
  b->n_h.assign(n, n + nb);
  b->xoff_h.resize(nb);
  long tot = 0, tch = 0;
  std::vector<long> choff(nb + 1);
  for (int i = 0; i < nb; ++i) {
    if (n[i] < k || n[i] >= KM_BLOCK) return KMEANS_ERROR_INVALID_ARGUMENT;
    b->xoff_h[i] = tot;
    tot += n[i];
    choff[i] = tch;
    tch += (n[i] + KM_CHUNK - 1) / KM_CHUNK + k;
  }
  choff[nb] = tch;
  if (tot > 0x7fffffffL) return KMEANS_ERROR_INVALID_ARGUMENT;
  b->total = tot;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    fprintf(stderr, "kmeans_amd: no HIP device available\n");
    return KMEANS_ERROR_BUILDDATA;
  }
  BCHK(hipSetDevice(b->opt.device), KMEANS_ERROR_BUILDDATA);
  if (b->opt.stream) {
    b->stream = (hipStream_t)b->opt.stream;
  } else {
    BCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), KMEANS_ERROR_BUILDDATA);
    b->own_stream = true;
  }
  std::vector<int> nv(nb), ord(nb);
  for (int i = 0; i < nb; ++i) nv[i] = (int)n[i];
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return n[a] > n[c]; });
  BCHK(hipMalloc(&b->xoff, sizeof(long) * nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->nvec, sizeof(int) * nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->order, sizeof(int) * nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->choff, sizeof(long) * (nb + 1)), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->desc, sizeof(int4) * (size_t)tch), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemcpy(b->xoff, b->xoff_h.data(), sizeof(long) * nb, hipMemcpyHostToDevice), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemcpy(b->nvec, nv.data(), sizeof(int) * nb, hipMemcpyHostToDevice), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemcpy(b->order, ord.data(), sizeof(int) * nb, hipMemcpyHostToDevice), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemcpy(b->choff, choff.data(), sizeof(long) * (nb + 1), hipMemcpyHostToDevice), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->C, sizeof(float) * (size_t)nb * k * d), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->nc, sizeof(float) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->bnd, sizeof(float2) * (size_t)nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->labels, sizeof(int) * (size_t)tot), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemsetAsync(b->labels, 0, sizeof(int) * (size_t)tot, b->stream), KMEANS_ERROR_BUILDDATA);
  for (int q = 0; q < 2; ++q) {
    // 1024 ints past the points: the staged sort's stores of rows past an
    // instance's end
    BCHK(hipMalloc(&b->perm[q], sizeof(int) * ((size_t)tot + 1024)), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMalloc(&b->part[q], sizeof(double) * (size_t)tch * (d + 1)), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMalloc(&b->gcnt[q], sizeof(int) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMalloc(&b->goff[q], sizeof(int) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMalloc(&b->gcpre[q], sizeof(int) * (size_t)nb * (k + 1)), KMEANS_ERROR_BUILDDATA);
  }
  BCHK(hipMalloc(&b->fd, sizeof(int) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->rstate, sizeof(int) * 2), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemset(b->rstate, 0, sizeof(int) * 2), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->imin, sizeof(int) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemset(b->imin, 0x7f, sizeof(int) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
  b->slots = tch;
  BCHK(hipMalloc(&b->counters, sizeof(int) * 2), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->nexact, sizeof(unsigned long long)), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemset(b->nexact, 0, sizeof(unsigned long long)), KMEANS_ERROR_BUILDDATA);
  // drift arrays (every engine's finalize writes them) and, for BF16 engines
  // whose centres fit the assign's LDS, the bounds and the open lists
  BCHK(hipMalloc(&b->drift, sizeof(float) * (size_t)nb * k), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->dmaxv, sizeof(float) * (size_t)nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->xmax, sizeof(float) * (size_t)nb), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->bvalid, sizeof(int)), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemset(b->bvalid, 0, sizeof(int)), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMalloc(&b->nkp, sizeof(unsigned long long) * 3), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemset(b->nkp, 0, sizeof(unsigned long long) * 3), KMEANS_ERROR_BUILDDATA);
  if (b->opt.precision == KM_PREC_BF16 && bf16_lds(b) <= 64 * 1024) {
    // the open lists of work items of any tile size P <= B_MAXP fill at most
    // tot + nb * P positions (their counts: upload_items)
    BCHK(hipMalloc(&b->xb, sizeof(float4) * (size_t)tot), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMalloc(&b->olist, sizeof(uint16_t) * ((size_t)tot + (size_t)nb * B_MAXP)),
         KMEANS_ERROR_BUILDDATA);
  }
  return KMEANS_SUCCESS;
}

static kmeansErrorStatus ensure_slots(Batch *b, int iters) {
  if (iters <= b->comp_slots) return KMEANS_SUCCESS;
  hipFree(b->comp);
  hipFree(b->empty);
  b->comp = nullptr;
  b->empty = nullptr;
  b->comp_slots = 0;
  BCHK(hipMalloc(&b->comp, sizeof(double) * (size_t)iters * b->nb), KMEANS_ERROR_COMPUTE);
  BCHK(hipMalloc(&b->empty, sizeof(int) * (size_t)iters * b->nb), KMEANS_ERROR_COMPUTE);
  b->comp_slots = iters;
  batch_free_graph(b);  // slot addresses are baked into the graph
  return KMEANS_SUCCESS;
}

// timed (kmeansOptions.timing): events before the assign, between assign and
// update and after the update of every iteration (an eager run, no graph)
static hipError_t enqueue_iterations(Batch *b, int iters, bool timed) {
  hipError_t e;
  while (timed && b->gev.size() < (size_t)3 * iters) {
    hipEvent_t ev = nullptr;
    if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
    b->gev.push_back(ev);
  }
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(64), 0, b->stream, b->counters, 2, 0);
    if (timed && (e = hipEventRecord(b->gev[3 * it], b->stream)) != hipSuccess) return e;
    if ((e = launch_assign(b, b->P, b->nitems, b->items)) != hipSuccess) return e;
    if (timed && (e = hipEventRecord(b->gev[3 * it + 1], b->stream)) != hipSuccess) return e;
    if ((e = launch_update(b, it)) != hipSuccess) return e;
    if (timed && (e = hipEventRecord(b->gev[3 * it + 2], b->stream)) != hipSuccess) return e;
  }
  return hipGetLastError();
}

static kmeansErrorStatus batch_run(Batch *b, int iters) {
  if (!b->data_set || !b->centers_set) return KMEANS_ERROR_COMPUTE;
  if (iters <= 0) return KMEANS_SUCCESS;
  kmeansErrorStatus st;
  if ((st = ensure_slots(b, iters)) != KMEANS_SUCCESS) return st;
  if (!b->P) {
    const int P = autotune(b);
    BCHK(upload_items(b, P), KMEANS_ERROR_COMPUTE);
  }
  const auto t0 = std::chrono::steady_clock::now();
  const bool timed = b->opt.timing != 0;
  if (!timed && b->graph_iters != iters) {
    batch_free_graph(b);
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), KMEANS_ERROR_COMPUTE);
    hipError_t e = enqueue_iterations(b, iters, false);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) hipGraphDestroy(g);
      fprintf(stderr, "kmeans_amd: batch graph capture failed: %s\n",
              hipGetErrorString(e != hipSuccess ? e : e2));
      return KMEANS_ERROR_COMPUTE;
    }
    b->graph = g;
    BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), KMEANS_ERROR_COMPUTE);
    b->graph_iters = iters;
  }
  const auto t1 = std::chrono::steady_clock::now();
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, b->stream);
  if (timed)
    BCHK(enqueue_iterations(b, iters, true), KMEANS_ERROR_COMPUTE);
  else
    BCHK(hipGraphLaunch(b->exec, b->stream), KMEANS_ERROR_COMPUTE);
  hipEventRecord(e1, b->stream);
  BCHK(hipEventSynchronize(e1), KMEANS_ERROR_COMPUTE);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  std::vector<int> emp((size_t)iters * b->nb);
  BCHK(hipMemcpy(emp.data(), b->empty, sizeof(int) * emp.size(), hipMemcpyDeviceToHost), KMEANS_ERROR_COMPUTE);
  long long es = 0;
  for (int v : emp) es += v;
  b->iters += iters;
  b->stats.iterations = b->iters;
  b->stats.ms_graph = ms;
  b->stats.ms_capture = std::chrono::duration<double, std::milli>(t1 - t0).count();
  b->stats.empty_clusters += es;
  b->stats.tile_rows = b->P;
  b->stats.work_items = b->nitems;
  size_t nnodes = 0;
  b->stats.graph_nodes = 0;
  if (!timed && hipGraphGetNodes(b->graph, nullptr, &nnodes) == hipSuccess)
    b->stats.graph_nodes = (int)nnodes;
  b->stats.ms_assign = b->stats.ms_update = 0.0;
  for (int it = 0; timed && it < iters; ++it) {
    float ta = 0.0f, tu = 0.0f;
    if (hipEventElapsedTime(&ta, b->gev[3 * it], b->gev[3 * it + 1]) == hipSuccess)
      b->stats.ms_assign += ta;
    if (hipEventElapsedTime(&tu, b->gev[3 * it + 1], b->gev[3 * it + 2]) == hipSuccess)
      b->stats.ms_update += tu;
  }
  b->stats.last_run_iters = iters;
  unsigned long long nx = 0, kp[3] = {0, 0, 0};
  BCHK(hipMemcpy(&nx, b->nexact, sizeof(nx), hipMemcpyDeviceToHost), KMEANS_ERROR_COMPUTE);
  BCHK(hipMemcpy(kp, b->nkp, sizeof(kp), hipMemcpyDeviceToHost), KMEANS_ERROR_COMPUTE);
  b->stats.exact_searched = (long long)nx;
  b->stats.kept = (long long)kp[0];
  b->stats.paired = (long long)kp[1];
  b->stats.scored = (long long)kp[2];
  return KMEANS_SUCCESS;
}

}  // namespace km

struct kmeansBatchEngine {
  km::Batch b;
};

using km::Batch;

extern "C" {

kmeansErrorStatus kmeansBatchCreate(kmeansBatchHandle *h, int instances, const long *n, int d,
                                    int k, const kmeansOptions *opt) {
  if (!h) return KMEANS_ERROR_INVALID_ARGUMENT;
  *h = nullptr;
  kmeansBatchEngine *p = new (std::nothrow) kmeansBatchEngine();
  if (!p) return KMEANS_ERROR_BUILDDATA;
  kmeansErrorStatus st = km::batch_create(&p->b, instances, n, d, k, opt);
  if (st != KMEANS_SUCCESS) {
    km::batch_free(&p->b);
    delete p;
    return st;
  }
  *h = p;
  return KMEANS_SUCCESS;
}

void kmeansBatchDestroy(kmeansBatchHandle h) {
  if (!h) return;
  km::batch_free(&h->b);
  delete h;
}

kmeansErrorStatus kmeansBatchSetData(kmeansBatchHandle h, const float *X, int on_device) {
  if (!h || !X) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  const size_t bytes = sizeof(float) * (size_t)b->total * b->d;
  if (on_device) {
    // the kernels read rows as 16-byte pieces when d is a multiple of 4
    if (b->d % 4 == 0 && ((uintptr_t)X & 15) != 0) {
      fprintf(stderr, "kmeans_amd: kmeansBatchSetData: borrowed device points must be 16-byte "
                      "aligned when d is a multiple of 4\n");
      return KMEANS_ERROR_INVALID_ARGUMENT;
    }
    b->X = X;
  } else {
    if (!b->X_owned) BCHK(hipMalloc(&b->X_owned, bytes), KMEANS_ERROR_BUILDDATA);
    BCHK(hipMemcpy(b->X_owned, X, bytes, hipMemcpyHostToDevice), KMEANS_ERROR_BUILDDATA);
    b->X = b->X_owned;
  }
  // new points: no drift bounds and no partials to reuse; their instances'
  // largest |x|
  BCHK(hipMemsetAsync(b->bvalid, 0, sizeof(int), b->stream), KMEANS_ERROR_BUILDDATA);
  BCHK(hipMemsetAsync(b->rstate, 0, sizeof(int) * 2, b->stream), KMEANS_ERROR_BUILDDATA);
  hipLaunchKernelGGL(km::k_batch_xmax, dim3(b->nb), dim3(256), 0, b->stream, b->d, b->xoff, b->nvec,
                     b->X, b->xmax);
  BCHK(hipStreamSynchronize(b->stream), KMEANS_ERROR_BUILDDATA);
  b->data_set = true;
  km::batch_free_graph(b);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchSetCenters(kmeansBatchHandle h, const float *C, int on_device) {
  if (!h || !C) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  const size_t nkd = (size_t)b->nb * b->k * b->d;
  BCHK(hipMemcpyAsync(b->C, C, sizeof(float) * nkd,
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, b->stream),
       KMEANS_ERROR_INITIALIZE);
  hipLaunchKernelGGL(km::k_batch_norms, dim3(km::bg1((long)b->nb * b->k)), dim3(256), 0, b->stream,
                     (long)b->nb * b->k, b->d, b->C, b->nc);
  hipLaunchKernelGGL(km::k_batch_bounds, dim3(b->nb), dim3(256), 0, b->stream, b->d, b->k, b->C,
                     b->nc, b->bnd);
  // centres set from outside: the drift bounds no longer hold
  BCHK(hipMemsetAsync(b->bvalid, 0, sizeof(int), b->stream), KMEANS_ERROR_INITIALIZE);
  BCHK(hipStreamSynchronize(b->stream), KMEANS_ERROR_INITIALIZE);
  b->centers_set = true;
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchInitFirstK(kmeansBatchHandle h) {
  if (!h || !h->b.data_set) return KMEANS_ERROR_INITIALIZE;
  Batch *b = &h->b;
  const long nkd = (long)b->nb * b->k * b->d;
  hipLaunchKernelGGL(km::k_batch_first_k, dim3(km::bg1(nkd)), dim3(256), 0, b->stream, b->nb, b->k,
                     b->d, b->xoff, b->X, b->C);
  hipLaunchKernelGGL(km::k_batch_norms, dim3(km::bg1((long)b->nb * b->k)), dim3(256), 0, b->stream,
                     (long)b->nb * b->k, b->d, b->C, b->nc);
  hipLaunchKernelGGL(km::k_batch_bounds, dim3(b->nb), dim3(256), 0, b->stream, b->d, b->k, b->C,
                     b->nc, b->bnd);
  // centres set from outside: the drift bounds no longer hold
  BCHK(hipMemsetAsync(b->bvalid, 0, sizeof(int), b->stream), KMEANS_ERROR_INITIALIZE);
  BCHK(hipStreamSynchronize(b->stream), KMEANS_ERROR_INITIALIZE);
  b->centers_set = true;
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchSetTileRows(kmeansBatchHandle h, int rows) {
  if (!h || rows < 0 || (rows && rows % 2) || rows > B_MAXP) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  km::batch_free_graph(b);
  if (rows == 0) {
    b->P = 0;
    return KMEANS_SUCCESS;
  }
  BCHK(km::upload_items(b, rows), KMEANS_ERROR_INVALID_ARGUMENT);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchRun(kmeansBatchHandle h, int iters) {
  if (!h || iters < 0) return KMEANS_ERROR_INVALID_ARGUMENT;
  return km::batch_run(&h->b, iters);
}

kmeansErrorStatus kmeansBatchGetCenters(kmeansBatchHandle h, float *C) {
  if (!h || !C) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  BCHK(hipMemcpy(C, b->C, sizeof(float) * (size_t)b->nb * b->k * b->d, hipMemcpyDeviceToHost),
       KMEANS_ERROR_GET_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchGetLabels(kmeansBatchHandle h, int *labels) {
  if (!h || !labels) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  BCHK(hipMemcpy(labels, b->labels, sizeof(int) * (size_t)b->total, hipMemcpyDeviceToHost),
       KMEANS_ERROR_GET_CENTERS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchGetCompactness(kmeansBatchHandle h, double *comp) {
  if (!h || !comp) return KMEANS_ERROR_INVALID_ARGUMENT;
  Batch *b = &h->b;
  const int it = b->stats.last_run_iters;
  if (it <= 0) return KMEANS_ERROR_COMPACTNESS;
  BCHK(hipMemcpy(comp, b->comp, sizeof(double) * (size_t)it * b->nb, hipMemcpyDeviceToHost),
       KMEANS_ERROR_COMPACTNESS);
  return KMEANS_SUCCESS;
}

kmeansErrorStatus kmeansBatchGetStats(kmeansBatchHandle h, kmeansBatchStats *s) {
  if (!h || !s) return KMEANS_ERROR_INVALID_ARGUMENT;
  *s = h->b.stats;
  return KMEANS_SUCCESS;
}

const float *kmeansBatchDeviceCenters(kmeansBatchHandle h) { return h ? h->b.C : nullptr; }
const int *kmeansBatchDeviceLabels(kmeansBatchHandle h) { return h ? h->b.labels : nullptr; }

}  // extern "C"
