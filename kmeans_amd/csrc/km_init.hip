// k-means++ seeding on the GPU (Arthur & Vassilvitskii, "k-means++: the
// advantages of careful seeding", SODA 2007).  The reference has no such
// initialisation: its initAlgorithm argument is ignored and the centres are
// rand() % m picks (kmeans.cpp:876-888).
//
// Canonical form (the CPU oracle's orc_kmeanspp computes the same picks):
//   dist(x, c) = left-to-right FP64 sum over d of ((double)x_d - (double)c_d)^2,
//                each difference and square rounded (no fma);
//   i_0 = rng_u64(seed, 6, 0) % m, D_i = dist(x_i, x_{i_0});
//   step t: W = the CSUM tree of D (level 0: left-to-right sums of 256
//   consecutive D_i, each further level the same over 256 consecutive sums
//   of the level below, up to a level of at most 256 values), T = the
//   left-to-right sum of the top level; T == 0: i_t = rng_u64(seed, 6, t) % m;
//   else target = rng_unif(seed, 6, t) * T and the tree is descended keeping a
//   running sum r (the prefix before the current node): the first child that
//   takes r past target is entered (none: the last child with a positive
//   value); D_i = min(D_i, dist(x_i, x_{i_t})).
//
// Per step: k_pp_dist (one block per 256 points: rows staged row-contiguously
// through LDS, the new distances and the level-0 sum of the block) and
// k_pp_pick (one block: the upper levels, the total and the descent; the
// pick stays in device memory, so the k - 1 steps run without host
// synchronisation).
#include <algorithm>
#include <vector>

#include "km_common.hpp"

namespace km {

template <typename T>
__global__ __launch_bounds__(256) void k_pp_dist(long m, int d, const T *__restrict__ X,
                                                 const int *__restrict__ idx, int t,
                                                 double *__restrict__ D, double *__restrict__ lv0) {
  __shared__ T tile[256][17];
  __shared__ T crow[16];
  __shared__ double part[256];
  const long r0 = (long)blockIdx.x * 256;
  const int rows = (int)min(256L, m - r0);
  const T *c = X + (long)idx[t] * d;
  double s = 0.0;
  for (int d0 = 0; d0 < d; d0 += 16) {
    const int dn = min(16, d - d0);
    __syncthreads();
    for (int e = threadIdx.x; e < rows * 16; e += 256) {
      const int r = e >> 4, q = e & 15;
      if (q < dn) tile[r][q] = X[(r0 + r) * d + d0 + q];
    }
    if (threadIdx.x < 16 && (int)threadIdx.x < dn) crow[threadIdx.x] = c[d0 + threadIdx.x];
    __syncthreads();
    if ((int)threadIdx.x < rows)
      for (int q = 0; q < dn; ++q) {
        const double v = (double)tile[threadIdx.x][q] - (double)crow[q];
        s += v * v;
      }
  }
  double nd = 0.0;
  if ((int)threadIdx.x < rows) {
    const long i = r0 + threadIdx.x;
    nd = t == 0 ? s : (s < D[i] ? s : D[i]);
    D[i] = nd;
  }
  part[threadIdx.x] = nd;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int r = 0; r < rows; ++r) a += part[r];
    lv0[blockIdx.x] = a;
  }
}

struct PPLevels {
  double *lv[6];  // lv[0]: one sum per 256 points, lv[l+1]: one per 256 of lv[l]
  long n[6];
  int L;          // number of levels; lv[L-1] has at most 256 values
};

// running sum over children v[0..n) from *r: the first child that takes it
// past target, else the last positive child; *r becomes the prefix before it
__device__ long pp_descend(const double *v, long n, double *r, double target) {
  double run = *r, last_before = *r;
  long last = -1;
  for (long i = 0; i < n; ++i) {
    const double before = run;
    run += v[i];
    if (v[i] > 0.0) { last = i; last_before = before; }
    if (run > target) { *r = before; return i; }
  }
  *r = last_before;
  return last < 0 ? n - 1 : last;
}

__global__ __launch_bounds__(256) void k_pp_pick(long m, const double *__restrict__ D,
                                                 PPLevels lv, uint64_t seed, int t,
                                                 int *__restrict__ idx) {
  // upper levels, each thread summing its 256 children left to right
  for (int l = 1; l < lv.L; ++l) {
    for (long g = threadIdx.x; g < lv.n[l]; g += blockDim.x) {
      const long a = g * KM_CHUNK, e = min(a + KM_CHUNK, lv.n[l - 1]);
      double s = 0.0;
      for (long i = a; i < e; ++i) s += lv.lv[l - 1][i];
      lv.lv[l][g] = s;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double *top = lv.lv[lv.L - 1];
  const long ntop = lv.n[lv.L - 1];
  double T = 0.0;
  for (long i = 0; i < ntop; ++i) T += top[i];
  long pick;
  if (T == 0.0) {
    pick = (long)(rng_u64(seed, 6, (uint64_t)t) % (uint64_t)m);
  } else {
    const double target = rng_unif(seed, 6, (uint64_t)t) * T;
    double r = 0.0;
    long node = pp_descend(top, ntop, &r, target);
    for (int l = lv.L - 1; l >= 0; --l) {
      const double *ch = l > 0 ? lv.lv[l - 1] : D;
      const long cn = l > 0 ? lv.n[l - 1] : m;
      const long a = node * KM_CHUNK, e = min(a + KM_CHUNK, cn);
      node = a + pp_descend(ch + a, e - a, &r, target);
    }
    pick = node;
  }
  idx[t] = (int)pick;
}

// k picked row indices of X (m x d) into idx (device, k ints); temporary
// device memory: m + about m/255 doubles.
template <typename T>
static hipError_t kmeanspp_t(hipStream_t s, long m, int d, int k, const T *X, uint64_t seed,
                             int *idx) {
  if (k <= 0 || k > m) return hipErrorInvalidValue;
  PPLevels lv{};
  long n = m, total = 0;
  int L = 0;
  do {
    n = km_cdiv(n, KM_CHUNK);
    lv.n[L++] = n;
    total += n;
  } while (n > KM_CHUNK && L < 6);
  if (n > KM_CHUNK) return hipErrorInvalidValue;
  lv.L = L;
  double *buf = nullptr;
  hipError_t e = hipMalloc(&buf, sizeof(double) * (size_t)(m + total));
  if (e != hipSuccess) return e;
  double *D = buf;
  long off = m;
  for (int l = 0; l < L; ++l) {
    lv.lv[l] = buf + off;
    off += lv.n[l];
  }
  const int i0 = (int)(rng_u64(seed, 6, 0) % (uint64_t)m);
  e = hipMemcpyAsync(idx, &i0, sizeof(int), hipMemcpyHostToDevice, s);
  const unsigned nb = (unsigned)km_cdiv(m, 256);
  for (int t = 0; t < k && e == hipSuccess; ++t) {
    if (t > 0) hipLaunchKernelGGL(k_pp_pick, dim3(1), dim3(256), 0, s, m, D, lv, seed, t, idx);
    hipLaunchKernelGGL(k_pp_dist<T>, dim3(nb), dim3(256), 0, s, m, d, X, idx, t, D, lv.lv[0]);
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(s);  // the host copy of i0 and buf must outlive the work
  hipFree(buf);
  return e != hipSuccess ? e : e2;
}

hipError_t kmeanspp(hipStream_t s, long m, int d, int k, const float *X, uint64_t seed, int *idx) {
  return kmeanspp_t(s, m, d, k, X, seed, idx);
}
hipError_t kmeanspp(hipStream_t s, long m, int d, int k, const double *X, uint64_t seed,
                    int *idx) {
  return kmeanspp_t(s, m, d, k, X, seed, idx);
}

}  // namespace km
