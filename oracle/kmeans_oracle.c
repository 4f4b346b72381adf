/*
 * kmeans_oracle.c -- CPU restatement of the Lloyd K-means hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (kmeans_amd/) never links, loads or calls it.
 *
 * What it restates (reference paths relative to the reference repo):
 *   orc_col_norms    main/cuda/Normalize.cu:33-47 (_dev_colNormalize), with
 *                    the square-accumulate written as an explicit fma chain.
 *   orc_assign       main/cuda/ClosestCenters.cu:55-69 (score starts at
 *                    -0.5f*||c||^2), multiply.hcu:27-53 + dotproducts.hcu:11-25
 *                    + fma.hcu:13 (fmaf over d in ascending order),
 *                    vectorMin.hcu:19-38 + reduction.hcu:60-65 +
 *                    ClosestCenters.cu:111-121 (lowest index among maxima).
 *   orc_update       main/cuda/ClusterCenters.cu:3-87 (count and mean of the
 *                    members of each centre) with the summation order fixed
 *                    to the blocked canonical order below and carried in FP64.
 *   orc_compactness  main/cuda/Compactness.cu:3-59 and main/cpp/kmeans.cpp:
 *                    1045-1104 (sum of squared distances to the NEW centres).
 *   orc_lloyd        main/cpp/kmeans.cpp:1186-1247 (loop, relErr criterion).
 *   orc_*_f64        the same in double precision (computeKmeansD,
 *                    kmeans.cpp:1364-1409, and the ...D kernel names,
 *                    KmeansCudaKernels.h:247-474): every float chain above
 *                    becomes the double chain (fma in double, scores start
 *                    at -0.5*nc_j), sums and moments are unchanged (they are
 *                    FP64 already), centres are S/n without the cast to
 *                    float.  The squares in T_c are the rounded x*x, which
 *                    for float data equals the FP32 definition below (those
 *                    squares are exact).
 *   orc_row_transform_min{f,d}  main/cuda/RowTransformMinimum.cu:3-58.
 *   orc_kmeanspp     k-means++ seeding (Arthur & Vassilvitskii 2007) in this
 *                    build's canonical form (see the function); the
 *                    reference has none (initAlgorithm is ignored).
 *
 * Canonical summation (the contract the GPU path reproduces bit for bit):
 *   CSUM(v[0..n)) = left-to-right FP64 sum from +0.0 if n <= KM_CHUNK,
 *                   else CSUM of the left-to-right sums of consecutive
 *                   KM_CHUNK-sized pieces (the last piece may be short).
 *   Points are cut into blocks of KM_BLOCK consecutive indices.
 *   Centre sum   S[j][d] = (((0 + B_0[j][d]) + B_1[j][d]) + ...) where
 *                B_b[j][d] = CSUM(x[i][d] : i in block b, label i == j, i asc).
 *   Moments      Q[j] = (((0 + QB_0[j]) + QB_1[j]) + ...).  The members of j in
 *                block b (i asc) are cut into KM_CHUNK pieces as CSUM cuts them;
 *                piece c gives q_c = left-to-right sum over d of T_c[d], T_c[d] =
 *                left-to-right FP64 sum over the piece of (double)x[i][d]^2 (each
 *                square exact in FP64); QB_b[j] = q_0 for one piece, else CSUM of
 *                the q_c, as CSUM combines piece sums.
 *   Compactness  comp = CSUM over j of comp_j, comp_j = (Q[j] - 2 a_j) + n_j b_j,
 *                a_j = left-to-right FP64 sum over d of (double)c'[j][d]*S[j][d],
 *                b_j = left-to-right FP64 sum over d of (double)c'[j][d]^2; this is
 *                sum_i |x_i - c'_{l_i}|^2 written with per-centre moments, so no
 *                second pass over the points is needed.
 *   New centre   c'[j][d] = (float)(S[j][d] / (double)n_j) if n_j > 0, else
 *                the previous centre (the reference divides 0/0 -> NaN,
 *                ClusterCenters.cu:81; this build keeps the old centre).
 *
 * Build: see oracle/Makefile (-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define KM_CHUNK 256
#define KM_BLOCK (1L << 23)

/* ------------------------------------------------------------------ */
/* Counter-based random numbers (SplitMix64 finaliser, Steele et al.,
 * "Fast splittable pseudorandom number generators", OOPSLA 2014).
 * u(seed, stream, i) is a pure function, so any element of any config
 * can be generated independently on the CPU or the GPU.               */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t k = mix64(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL);
  return mix64(k + i * 0x9E3779B97F4A7C15ULL);
}
/* uniform in (0,1): 53-bit mantissa, never 0 */
static inline double rng_unif(uint64_t seed, uint64_t stream, uint64_t i) {
  return ((double)(rng_u64(seed, stream, i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
/* standard normal by Box-Muller (cosine branch) */
static inline double rng_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = rng_unif(seed, stream, 2 * i), u2 = rng_unif(seed, stream, 2 * i + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t i) { return rng_u64(seed, stream, i); }
double orc_rng_unif(uint64_t seed, uint64_t stream, uint64_t i) { return rng_unif(seed, stream, i); }

/* Gaussian mixture: means U[mlo,mhi]^d, per-component sigma U[slo,shi],
 * component weights equal (zipf_s == 0) or Zipf(zipf_s) over ranks 1..k.
 * Streams: 0 means, 1 sigmas, 2 component choice, 3 noise.           */
void orc_gen_means(uint64_t seed, int d, int k, float mlo, float mhi, float slo, float shi,
                   float *means, float *sigmas) {
  for (long j = 0; j < (long)k * d; ++j)
    means[j] = (float)(mlo + (mhi - mlo) * rng_unif(seed, 0, (uint64_t)j));
  for (int j = 0; j < k; ++j)
    sigmas[j] = (float)(slo + (shi - slo) * rng_unif(seed, 1, (uint64_t)j));
}

/* cumulative weights for component choice (k entries, last == 1) */
void orc_gen_cdf(int k, double zipf_s, double *cdf) {
  double tot = 0.0;
  for (int j = 0; j < k; ++j) tot += (zipf_s > 0.0) ? pow((double)(j + 1), -zipf_s) : 1.0;
  double acc = 0.0;
  for (int j = 0; j < k; ++j) {
    acc += ((zipf_s > 0.0) ? pow((double)(j + 1), -zipf_s) : 1.0) / tot;
    cdf[j] = acc;
  }
  cdf[k - 1] = 1.0;
}

static inline int pick_component(const double *cdf, int k, double u) {
  int lo = 0, hi = k - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (u <= cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

/* rows [i0, i0+n) of the mixture into X (n x d row-major) */
void orc_gen_points(uint64_t seed, long i0, long n, int d, int k, const float *means,
                    const float *sigmas, const double *cdf, float *X, int *z) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n; ++r) {
    long i = i0 + r;
    int c = pick_component(cdf, k, rng_unif(seed, 2, (uint64_t)i));
    if (z) z[r] = c;
    for (int q = 0; q < d; ++q) {
      double g = rng_normal(seed, 3, (uint64_t)i * (uint64_t)d + (uint64_t)q);
      X[r * d + q] = (float)((double)means[(long)c * d + q] + (double)sigmas[c] * g);
    }
  }
}

/* k distinct indices in [0,n) by Floyd's algorithm, returned in draw order */
int orc_sample_init(uint64_t seed, long n, int k, int *idx) {
  if (k > n) return -1;
  long t = 0;
  for (long j = n - k; j < n; ++j) {
    long r = (long)(rng_u64(seed, 4, (uint64_t)t) % (uint64_t)(j + 1));
    int dup = 0;
    for (long q = 0; q < t; ++q) if (idx[q] == r) { dup = 1; break; }
    idx[t++] = dup ? (int)j : (int)r;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* rowNormalizeF (Normalize.cu:3-31, KmeansCudaKernels.h:428-432): out[i] =
 * fma chain over d ascending of x_id*x_id, the same chain as the centre
 * norms (the reference adds the squares in per-thread partial sums and a
 * block tree; this build fixes one order, the sequential one). */
void orc_row_norms(const float *X, long n, int d, float *out) {
  for (long i = 0; i < n; ++i) {
    float s = 0.0f;
    for (int q = 0; q < d; ++q) s = fmaf(X[i * d + q], X[i * d + q], s);
    out[i] = s;
  }
}

/* Centre norms: nc_j = fma chain over d of c_jd*c_jd (Normalize.cu:41-44). */
void orc_col_norms(const float *C, int k, int d, float *nc) {
  for (int j = 0; j < k; ++j) {
    float s = 0.0f;
    for (int q = 0; q < d; ++q) s = fmaf(C[(long)j * d + q], C[(long)j * d + q], s);
    nc[j] = s;
  }
}

/* Score and label.  s_ij = fma chain over d from -0.5f*nc_j; label = lowest
 * j among the maxima (strict '>' while scanning j ascending).  best/second
 * (optional) receive the largest and second-largest score over j.
 * SIMD runs across centres; each s_ij is still one sequential chain.  */
#define PB 8
#define JB 256
void orc_assign(const float *X, long n, int d, const float *C, int k, const float *nc,
                int *labels, float *best_out, float *second_out) {
  float *Ct = (float *)malloc(sizeof(float) * (size_t)d * (size_t)k);
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < d; ++q) Ct[(long)q * k + j] = C[(long)j * d + q];
  long nb = (n + PB - 1) / PB;
#pragma omp parallel
  {
    float s[PB][JB];
    float best[PB], second[PB];
    int lab[PB];
#pragma omp for schedule(dynamic, 16)
    for (long b = 0; b < nb; ++b) {
      long i0 = b * PB;
      int np = (int)((n - i0) < PB ? (n - i0) : PB);
      for (int p = 0; p < PB; ++p) { best[p] = -INFINITY; second[p] = -INFINITY; lab[p] = 0; }
      for (int j0 = 0; j0 < k; j0 += JB) {
        int nj = (k - j0) < JB ? (k - j0) : JB;
        for (int p = 0; p < np; ++p)
          for (int j = 0; j < nj; ++j) s[p][j] = -0.5f * nc[j0 + j];
        for (int q = 0; q < d; ++q) {
          const float *crow = Ct + (long)q * k + j0;
          for (int p = 0; p < np; ++p) {
            float xv = X[(i0 + p) * d + q];
            float *sp = s[p];
#pragma omp simd
            for (int j = 0; j < nj; ++j) sp[j] = fmaf(xv, crow[j], sp[j]);
          }
        }
        for (int p = 0; p < np; ++p)
          for (int j = 0; j < nj; ++j) {
            float v = s[p][j];
            if (v > best[p]) { second[p] = best[p]; best[p] = v; lab[p] = j0 + j; }
            else if (v > second[p]) second[p] = v;
          }
      }
      for (int p = 0; p < np; ++p) {
        labels[i0 + p] = lab[p];
        if (best_out) best_out[i0 + p] = best[p];
        if (second_out) second_out[i0 + p] = second[p];
      }
    }
  }
  free(Ct);
}

/* Score of one pair, same chain as orc_assign (for tests of single pairs). */
float orc_score(const float *x, const float *c, int d, float ncj) {
  float s = -0.5f * ncj;
  for (int q = 0; q < d; ++q) s = fmaf(x[q], c[q], s);
  return s;
}

/* ------------------------------------------------------------------ */
/* Canonical FP64 sum (see header). */
static double seq_sum(const double *v, long n) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += v[i];
  return s;
}
double orc_csum(const double *v, long n) {
  if (n <= KM_CHUNK) return seq_sum(v, n);
  long m = (n + KM_CHUNK - 1) / KM_CHUNK;
  double *p = (double *)malloc(sizeof(double) * (size_t)m);
  for (long c = 0; c < m; ++c) {
    long a = c * KM_CHUNK, e = a + KM_CHUNK < n ? a + KM_CHUNK : n;
    p[c] = seq_sum(v + a, e - a);
  }
  double r = orc_csum(p, m);
  free(p);
  return r;
}

/* The squared-norm moment of one chunk of members (rows[0..m) of X, m <=
 * KM_CHUNK, ascending point order): for each dimension q the left-to-right
 * FP64 sum over the members of x_q * x_q (exact for float x; for double x
 * each square is rounded once, no fma), then the left-to-right sum of those
 * d values in ascending q.  The per-dimension sums are what the chunk-sum
 * kernels accumulate beside the coordinate sums from the rows they read, so
 * no per-point squared norm is gathered. */
#define DEF_CHUNK_SQ(NAME, T)                                                              \
  static double NAME(const T *X, const long *rows, long m, int d) {                        \
    double tot = 0.0;                                                                      \
    for (int q = 0; q < d; ++q) {                                                          \
      double t = 0.0;                                                                      \
      for (long r = 0; r < m; ++r) {                                                       \
        const double v = (double)X[rows[r] * d + q];                                       \
        t += v * v;                                                                        \
      }                                                                                    \
      tot += t;                                                                            \
    }                                                                                      \
    return tot;                                                                            \
  }
DEF_CHUNK_SQ(chunk_sq, float)
DEF_CHUNK_SQ(chunk_sq_f64, double)

/* Per-block canonical centre sums.  For block b and each centre j, collect
 * the members in ascending i (stable counting sort by label), then CSUM each
 * dimension; the squared norms are summed per chunk of KM_CHUNK members
 * (chunk_sq) and the chunk values combined as CSUM combines chunk partials.
 * Adds B_b[j][d] into S (FP64, K x d), QB_b[j] into Q and n_b into counts. */
#define DEF_BLOCK_UPDATE(NAME, T, PSQ)                                                     \
  static void NAME(const T *X, long i0, long n, int d, const int *labels, int k,           \
                   double *S, double *Q, long long *counts) {                              \
    long *cnt = (long *)calloc((size_t)k + 1, sizeof(long));                               \
    for (long r = 0; r < n; ++r) cnt[labels[i0 + r] + 1]++;                                \
    for (int j = 0; j < k; ++j) cnt[j + 1] += cnt[j];                                      \
    long *off = (long *)malloc(sizeof(long) * ((size_t)k + 1));                            \
    memcpy(off, cnt, sizeof(long) * ((size_t)k + 1));                                      \
    long *perm = (long *)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));                   \
    for (long r = 0; r < n; ++r) perm[off[labels[i0 + r]]++] = i0 + r;                     \
    _Pragma("omp parallel")                                                                \
    {                                                                                      \
      double *v = NULL;                                                                    \
      long vcap = 0;                                                                       \
      _Pragma("omp for schedule(dynamic, 1)")                                              \
      for (int j = 0; j < k; ++j) {                                                        \
        long a = cnt[j], m = cnt[j + 1] - cnt[j];                                          \
        counts[j] += m;                                                                    \
        if (m > vcap) { free(v); vcap = m; v = (double *)malloc(sizeof(double) * (size_t)m); } \
        for (int q = 0; q < d; ++q) {                                                      \
          for (long t = 0; t < m; ++t) v[t] = (double)X[perm[a + t] * d + q];              \
          S[(long)j * d + q] += orc_csum(v, m);                                            \
        }                                                                                  \
        {                                                                                  \
          const long nq = (m + KM_CHUNK - 1) / KM_CHUNK;                                   \
          for (long c = 0; c < nq; ++c) {                                                  \
            const long ca = c * KM_CHUNK, cm = m - ca < KM_CHUNK ? m - ca : KM_CHUNK;      \
            v[c] = PSQ(X, perm + a + ca, cm, d);                                           \
          }                                                                                \
          if (m > 0) Q[j] += nq == 1 ? v[0] : orc_csum(v, nq);                             \
        }                                                                                  \
      }                                                                                    \
      free(v);                                                                             \
    }                                                                                      \
    free(perm); free(off); free(cnt);                                                      \
  }
DEF_BLOCK_UPDATE(block_update, float, chunk_sq)
DEF_BLOCK_UPDATE(block_update_f64, double, chunk_sq_f64)

/* Canonical per-centre moments over all points: S (k x d), Q (k), counts. */
#define DEF_MOMENTS(NAME, T, BLK)                                                          \
  void NAME(const T *X, long n, int d, const int *labels, int k, double *S, double *Q,     \
            long long *counts) {                                                           \
    memset(S, 0, sizeof(double) * (size_t)k * d);                                          \
    memset(Q, 0, sizeof(double) * (size_t)k);                                              \
    memset(counts, 0, sizeof(long long) * (size_t)k);                                      \
    for (long i0 = 0; i0 < n; i0 += KM_BLOCK) {                                            \
      long nb = (n - i0) < KM_BLOCK ? (n - i0) : KM_BLOCK;                                 \
      BLK(X, i0, nb, d, labels, k, S, Q, counts);                                          \
    }                                                                                      \
  }
DEF_MOMENTS(orc_moments, float, block_update)
DEF_MOMENTS(orc_moments_f64, double, block_update_f64)

/* Compactness from moments and centres C (k x d, any centres). */
#define DEF_COMP_MOMENTS(NAME, T)                                                          \
  double NAME(const double *S, const double *Q, const long long *counts, const T *C, int k,  \
              int d) {                                                                     \
    double *cj = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));               \
    for (int j = 0; j < k; ++j) {                                                          \
      double a = 0.0, b = 0.0;                                                             \
      for (int q = 0; q < d; ++q) {                                                        \
        const double c = (double)C[(long)j * d + q];                                       \
        a += c * S[(long)j * d + q];                                                       \
        b += c * c;                                                                        \
      }                                                                                    \
      cj[j] = (Q[j] - 2.0 * a) + (double)counts[j] * b;                                    \
    }                                                                                      \
    double r = orc_csum(cj, k);                                                            \
    free(cj);                                                                              \
    return r;                                                                              \
  }
DEF_COMP_MOMENTS(orc_compactness_moments_f64, double)
double orc_compactness_moments(const double *S, const double *Q, const long long *counts,
                               const float *C, int k, int d) {
  double *cj = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
  for (int j = 0; j < k; ++j) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < d; ++q) {
      const double c = (double)C[(long)j * d + q];
      a += c * S[(long)j * d + q];
      b += c * c;
    }
    cj[j] = (Q[j] - 2.0 * a) + (double)counts[j] * b;
  }
  double r = orc_csum(cj, k);
  free(cj);
  return r;
}

/* Update: new centres from labels (canonical order).  C_old/C_new are K x d
 * row-major; C_new may alias C_old.  counts (K), S (K x d) and Q (K) optional;
 * comp (optional) receives the compactness against the new centres. */
int orc_update(const float *X, long n, int d, const int *labels, int k, const float *C_old,
               float *C_new, long long *counts_out, double *S_out, double *Q_out, double *comp) {
  double *S = (double *)malloc(sizeof(double) * (size_t)k * d);
  double *Q = (double *)malloc(sizeof(double) * (size_t)k);
  long long *counts = (long long *)malloc(sizeof(long long) * (size_t)k);
  orc_moments(X, n, d, labels, k, S, Q, counts);
  int empty = 0;
  for (int j = 0; j < k; ++j) {
    if (counts[j] > 0) {
      for (int q = 0; q < d; ++q)
        C_new[(long)j * d + q] = (float)(S[(long)j * d + q] / (double)counts[j]);
    } else {
      ++empty;
      if (C_new != C_old)
        for (int q = 0; q < d; ++q) C_new[(long)j * d + q] = C_old[(long)j * d + q];
    }
  }
  if (comp) *comp = orc_compactness_moments(S, Q, counts, C_new, k, d);
  if (counts_out) memcpy(counts_out, counts, sizeof(long long) * (size_t)k);
  if (S_out) memcpy(S_out, S, sizeof(double) * (size_t)k * d);
  if (Q_out) memcpy(Q_out, Q, sizeof(double) * (size_t)k);
  free(S); free(Q); free(counts);
  return empty;
}

/* Compactness of labels against arbitrary centres C (k x d). */
double orc_compactness(const float *X, long n, int d, const int *labels, const float *C, int k) {
  double *S = (double *)malloc(sizeof(double) * (size_t)k * d);
  double *Q = (double *)malloc(sizeof(double) * (size_t)k);
  long long *counts = (long long *)malloc(sizeof(long long) * (size_t)k);
  orc_moments(X, n, d, labels, k, S, Q, counts);
  double r = orc_compactness_moments(S, Q, counts, C, k, d);
  free(S); free(Q); free(counts);
  return r;
}

/* ------------------------------------------------------------------ */
/* Full Lloyd loop (kmeans.cpp:1186-1247): while (it < maxIters && relErr >
 * criterion): assign, update, compactness against the new centres,
 * relErr = |comp - prev| / |comp| with prev = 0 before the first iteration.
 * C (K x d) holds C0 on entry and the final centres on exit.
 * Returns 0 on convergence (relErr <= criterion), 13 otherwise.        */
int orc_lloyd(const float *X, long n, int d, int k, float *C, int maxIters, double criterion,
              int *labels, double *comp_hist, int *iters_out, double *relerr_out) {
  float *nc = (float *)malloc(sizeof(float) * (size_t)k);
  double prev = 0.0, relErr = 1.0;
  int it = 0;
  while (it < maxIters && relErr > criterion) {
    orc_col_norms(C, k, d, nc);
    orc_assign(X, n, d, C, k, nc, labels, NULL, NULL);
    double comp = 0.0;
    orc_update(X, n, d, labels, k, C, C, NULL, NULL, NULL, &comp);
    relErr = fabs(comp - prev) / fabs(comp);
    prev = comp;
    if (comp_hist) comp_hist[it] = comp;
    ++it;
  }
  free(nc);
  if (iters_out) *iters_out = it;
  if (relerr_out) *relerr_out = relErr;
  return relErr <= criterion ? 0 : 13;
}

/* ------------------------------------------------------------------ */
/* k-means++ seeding (no counterpart in the reference, whose initAlgorithm is
 * ignored, kmeans.cpp:876-888; the seeding of Arthur & Vassilvitskii,
 * SODA 2007).  Canonical form:
 *   dist(x, c) = left-to-right FP64 sum over d of ((double)x_d - (double)c_d)^2
 *                (difference and square each rounded, no fma);
 *   i_0 = rng_u64(seed, 6, 0) % m;  D_i = dist(x_i, x_{i_0});
 *   step t = 1..k-1: W = the CSUM tree of D (level 0: left-to-right sums of
 *   256 consecutive D_i, level l+1: the same over 256 consecutive level-l
 *   sums, up to a level of at most 256 values); T = left-to-right sum of the
 *   top level; if T == 0, i_t = rng_u64(seed, 6, t) % m; otherwise with
 *   target = rng_unif(seed, 6, t) * T, descend the tree: at each level keep a
 *   running sum r (the prefix before the current node, starting from 0),
 *   add the node's children left to right and enter the first child that
 *   takes r past target (if none does, the last child with a positive
 *   value, r then being the running sum before it); at level 0 the child is
 *   the point.  D_i = min(D_i, dist(x_i, x_{i_t})).                       */
#define DEF_PP_DIST(NAME, T)                                                   \
  static double NAME(const T *x, const T *c, int d) {                          \
    double s = 0.0;                                                            \
    for (int q = 0; q < d; ++q) {                                              \
      const double v = (double)x[q] - (double)c[q];                            \
      s += v * v;                                                              \
    }                                                                          \
    return s;                                                                  \
  }
DEF_PP_DIST(pp_dist, float)
DEF_PP_DIST(pp_dist_f64, double)

/* descend one level: children v[0..n), running sum *r; returns the child */
static long pp_descend(const double *v, long n, double *r, double target) {
  double run = *r, before = *r;
  long last = -1;
  double last_before = *r;
  for (long i = 0; i < n; ++i) {
    before = run;
    run += v[i];
    if (v[i] > 0.0) { last = i; last_before = before; }
    if (run > target) { *r = before; return i; }
  }
  *r = last_before;
  return last < 0 ? n - 1 : last;
}

#define DEF_KMEANSPP(NAME, ET, DIST)                                                   \
int NAME(const ET *X, long m, int d, int k, uint64_t seed, int *idx) {                 \
  if (k > m || k <= 0) return -1;                                                      \
  double *D = (double *)malloc(sizeof(double) * (size_t)m);                            \
  /* tree levels: lv[0] = sums of 256 points, ... */                                   \
  double *lv[8];                                                                       \
  long ln[8];                                                                          \
  int L = 0;                                                                           \
  long n = m;                                                                          \
  while (n > KM_CHUNK) {                                                               \
    ln[L] = (n + KM_CHUNK - 1) / KM_CHUNK;                                             \
    lv[L] = (double *)malloc(sizeof(double) * (size_t)ln[L]);                          \
    n = ln[L];                                                                         \
    ++L;                                                                               \
  }                                                                                    \
  idx[0] = (int)(rng_u64(seed, 6, 0) % (uint64_t)m);                                   \
  {                                                                                    \
    const ET *c = X + (long)idx[0] * d;                                                \
_Pragma("omp parallel for schedule(static)")                                           \
    for (long i = 0; i < m; ++i) D[i] = DIST(X + i * d, c, d);                         \
  }                                                                                    \
  for (int t = 1; t < k; ++t) {                                                        \
    /* the tree */                                                                     \
    const double *below = D;                                                           \
    long nb = m;                                                                       \
    for (int l = 0; l < L; ++l) {                                                      \
      for (long g = 0; g < ln[l]; ++g) {                                               \
        const long a = g * KM_CHUNK, e = a + KM_CHUNK < nb ? a + KM_CHUNK : nb;        \
        double s = 0.0;                                                                \
        for (long i = a; i < e; ++i) s += below[i];                                    \
        lv[l][g] = s;                                                                  \
      }                                                                                \
      below = lv[l];                                                                   \
      nb = ln[l];                                                                      \
    }                                                                                  \
    double tot = 0.0;                                                                  \
    for (long i = 0; i < nb; ++i) tot += below[i];                                     \
    long pick;                                                                         \
    if (tot == 0.0) {                                                                  \
      pick = (long)(rng_u64(seed, 6, (uint64_t)t) % (uint64_t)m);                      \
    } else {                                                                           \
      const double target = rng_unif(seed, 6, (uint64_t)t) * tot;                      \
      double r = 0.0;                                                                  \
      long node = pp_descend(below, nb, &r, target);  /* top level */                  \
      for (int l = L - 1; l >= 0; --l) {                                               \
        const double *ch = l > 0 ? lv[l - 1] : D;                                      \
        const long cn = l > 0 ? ln[l - 1] : m;                                         \
        const long a = node * KM_CHUNK, e = a + KM_CHUNK < cn ? a + KM_CHUNK : cn;     \
        node = a + pp_descend(ch + a, e - a, &r, target);                              \
      }                                                                                \
      pick = node;                                                                     \
    }                                                                                  \
    idx[t] = (int)pick;                                                                \
    const ET *c = X + pick * d;                                                        \
_Pragma("omp parallel for schedule(static)")                                           \
    for (long i = 0; i < m; ++i) {                                                     \
      const double v = DIST(X + i * d, c, d);                                          \
      if (v < D[i]) D[i] = v;                                                          \
    }                                                                                  \
  }                                                                                    \
  for (int l = 0; l < L; ++l) free(lv[l]);                                             \
  free(D);                                                                             \
  return 0;                                                                            \
}
DEF_KMEANSPP(orc_kmeanspp, float, pp_dist)
DEF_KMEANSPP(orc_kmeanspp_f64, double, pp_dist_f64)

/* ------------------------------------------------------------------ */
/* Double-precision engine (computeKmeansD and the ...D kernel names). */
void orc_col_norms_f64(const double *C, int k, int d, double *nc) {
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
    for (int q = 0; q < d; ++q) s = fma(C[(long)j * d + q], C[(long)j * d + q], s);
    nc[j] = s;
  }
}

/* rowNormalizeD (Normalize.cu:3-31): fma chain of the squares of each row. */
void orc_row_norms_f64(const double *X, long n, int d, double *out) {
  for (long i = 0; i < n; ++i) {
    double s = 0.0;
    for (int q = 0; q < d; ++q) s = fma(X[i * d + q], X[i * d + q], s);
    out[i] = s;
  }
}

double orc_score_f64(const double *x, const double *c, int d, double ncj) {
  double s = -0.5 * ncj;
  for (int q = 0; q < d; ++q) s = fma(x[q], c[q], s);
  return s;
}

/* Labels in double: s_ij = fma chain from -0.5*nc_j, lowest index among the
 * maxima.  One point per iteration, centres in order (no SIMD needed for the
 * sizes the tests use). */
void orc_assign_f64(const double *X, long n, int d, const double *C, int k, const double *nc,
                    int *labels) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double best = -INFINITY;
    int lab = 0;
    for (int j = 0; j < k; ++j) {
      const double v = orc_score_f64(X + i * d, C + (long)j * d, d, nc[j]);
      if (v > best) { best = v; lab = j; }
    }
    labels[i] = lab;
  }
}

int orc_update_f64(const double *X, long n, int d, const int *labels, int k, const double *C_old,
                   double *C_new, long long *counts_out, double *comp) {
  double *S = (double *)malloc(sizeof(double) * (size_t)k * d);
  double *Q = (double *)malloc(sizeof(double) * (size_t)k);
  long long *counts = (long long *)malloc(sizeof(long long) * (size_t)k);
  orc_moments_f64(X, n, d, labels, k, S, Q, counts);
  int empty = 0;
  for (int j = 0; j < k; ++j) {
    if (counts[j] > 0) {
      for (int q = 0; q < d; ++q) C_new[(long)j * d + q] = S[(long)j * d + q] / (double)counts[j];
    } else {
      ++empty;
      if (C_new != C_old)
        for (int q = 0; q < d; ++q) C_new[(long)j * d + q] = C_old[(long)j * d + q];
    }
  }
  if (comp) *comp = orc_compactness_moments_f64(S, Q, counts, C_new, k, d);
  if (counts_out) memcpy(counts_out, counts, sizeof(long long) * (size_t)k);
  free(S); free(Q); free(counts);
  return empty;
}

double orc_compactness_f64(const double *X, long n, int d, const int *labels, const double *C,
                           int k) {
  double *S = (double *)malloc(sizeof(double) * (size_t)k * d);
  double *Q = (double *)malloc(sizeof(double) * (size_t)k);
  long long *counts = (long long *)malloc(sizeof(long long) * (size_t)k);
  orc_moments_f64(X, n, d, labels, k, S, Q, counts);
  double r = orc_compactness_moments_f64(S, Q, counts, C, k, d);
  free(S); free(Q); free(counts);
  return r;
}

/* Lloyd loop in double (same control flow as orc_lloyd). */
int orc_lloyd_f64(const double *X, long n, int d, int k, double *C, int maxIters,
                  double criterion, int *labels, double *comp_hist, int *iters_out,
                  double *relerr_out) {
  double *nc = (double *)malloc(sizeof(double) * (size_t)k);
  double prev = 0.0, relErr = 1.0;
  int it = 0;
  while (it < maxIters && relErr > criterion) {
    orc_col_norms_f64(C, k, d, nc);
    orc_assign_f64(X, n, d, C, k, nc, labels);
    double comp = 0.0;
    orc_update_f64(X, n, d, labels, k, C, C, NULL, &comp);
    relErr = fabs(comp - prev) / fabs(comp);
    prev = comp;
    if (comp_hist) comp_hist[it] = comp;
    ++it;
  }
  free(nc);
  if (iters_out) *iters_out = it;
  if (relerr_out) *relerr_out = relErr;
  return relErr <= criterion ? 0 : 13;
}

/* rowTransformMinimum (RowTransformMinimum.cu:3-58): for row i of the m x n
 * product P, v_ij = (T)((double)(rn[i] + cn[j]) - 2.0 * (double)P[i][j])
 * (the first sum in T: TYPE + TYPE, then a double expression from the
 * 2.0 literal, assigned back to TYPE); out[m0 + i] = the lowest j with the
 * smallest v below FLT_MAX (the start value for both types, :17), or -1 if
 * there is none.  rn is
 * indexed by the local row i (:20). */
void orc_row_transform_min_f(long m0, long m, int n, const float *rn, const float *cn,
                             const float *P, int *out) {
  for (long i = 0; i < m; ++i) {
    float bv = 3.40282346638528859812e+38f;
    int bi = -1;
    for (int j = 0; j < n; ++j) {
      const float v = (float)((double)(rn[i] + cn[j]) - 2.0 * (double)P[i * n + j]);
      if (v < bv) { bv = v; bi = j; }
    }
    out[m0 + i] = bi;
  }
}
void orc_row_transform_min_d(long m0, long m, int n, const double *rn, const double *cn,
                             const double *P, int *out) {
  for (long i = 0; i < m; ++i) {
    double bv = 3.40282346638528859812e+38;  /* FLT_MAX for both types (:17) */
    int bi = -1;
    for (int j = 0; j < n; ++j) {
      const double v = (rn[i] + cn[j]) - 2.0 * P[i * n + j];
      if (v < bv) { bv = v; bi = j; }
    }
    out[m0 + i] = bi;
  }
}

/* ------------------------------------------------------------------ */
/* FP64 oracle: exact-as-possible reference values for error reports.
 * Means by long-double sums, compactness in FP64 against given centres. */
void orc_means_fp64(const float *X, long n, int d, const int *labels, int k, double *M,
                    long long *counts) {
  long double *acc = (long double *)calloc((size_t)k * d, sizeof(long double));
  memset(counts, 0, sizeof(long long) * (size_t)k);
  for (long i = 0; i < n; ++i) {
    int j = labels[i];
    counts[j]++;
    for (int q = 0; q < d; ++q) acc[(long)j * d + q] += (long double)X[i * d + q];
  }
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < d; ++q)
      M[(long)j * d + q] = counts[j] ? (double)(acc[(long)j * d + q] / (long double)counts[j]) : 0.0;
  free(acc);
}

double orc_compactness_fp64(const float *X, long n, int d, const int *labels, const float *C) {
  long double tot = 0.0L;
#pragma omp parallel
  {
    long double loc = 0.0L;
#pragma omp for schedule(static)
    for (long i = 0; i < n; ++i) {
      const float *c = C + (long)labels[i] * d;
      double e = 0.0;
      for (int q = 0; q < d; ++q) {
        double v = (double)X[i * d + q] - (double)c[q];
        e += v * v;
      }
      loc += e;
    }
#pragma omp critical
    tot += loc;
  }
  return (double)tot;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { omp_set_num_threads(t); }

/* ------------------------------------------------------------------ */
/* Candidate-group proposer MLP (no reference counterpart: the learned
 * pruning of the assignment).  din -> h -> h -> g, ReLU, softmax
 * cross-entropy, mean over the batch.  Parameters as one flat vector:
 * W1 (h x din), b1 (h), W2 (h x h), b2 (h), W3 (g x h), b3 (g).  FP64
 * arithmetic; with mirror != 0 the operands of every product are rounded to
 * BF16 where the GPU rounds them (weights, hidden activations, the three
 * back-propagated signals), so the comparison isolates accumulation error. */
static double bf16r(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
  }
  return (double)f;
}

double orc_mlp_fwd_bwd(int din, int h, int g, int B, const float *X, const int *t, const double *P,
                       double *G, int mirror) {
  const double *W1 = P, *b1 = W1 + (long)h * din, *W2 = b1 + h, *b2 = W2 + (long)h * h;
  const double *W3 = b2 + h, *b3 = W3 + (long)g * h;
  double *gW1 = G, *gb1 = gW1 + (long)h * din, *gW2 = gb1 + h, *gb2 = gW2 + (long)h * h;
  double *gW3 = gb2 + h, *gb3 = gW3 + (long)g * h;
  const long P_n = (long)h * din + h + (long)h * h + h + (long)g * h + g;
  memset(G, 0, sizeof(double) * (size_t)P_n);
#define R(v) (mirror ? bf16r(v) : (v))
  double *w1 = malloc(sizeof(double) * (size_t)h * din), *w2 = malloc(sizeof(double) * (size_t)h * h);
  double *w3 = malloc(sizeof(double) * (size_t)g * h);
  for (long i = 0; i < (long)h * din; ++i) w1[i] = R(W1[i]);
  for (long i = 0; i < (long)h * h; ++i) w2[i] = R(W2[i]);
  for (long i = 0; i < (long)g * h; ++i) w3[i] = R(W3[i]);
  double *H1 = malloc(sizeof(double) * (size_t)B * h), *H2 = malloc(sizeof(double) * (size_t)B * h);
  double *dZ = malloc(sizeof(double) * (size_t)B * g), *dH = malloc(sizeof(double) * (size_t)B * h);
  double *dHr = malloc(sizeof(double) * (size_t)B * h), *dZr = malloc(sizeof(double) * (size_t)B * g);
  double loss = 0.0;
  double *z = calloc((size_t)g, sizeof(double));
  for (int r = 0; r < B; ++r) {
    const float *x = X + (long)r * din;
    for (int o = 0; o < h; ++o) {
      double a = 0.0;
      for (int i = 0; i < din; ++i) a += w1[(long)o * din + i] * (double)x[i];
      a += b1[o];
      H1[(long)r * h + o] = R(a > 0.0 ? a : 0.0);
    }
    for (int o = 0; o < h; ++o) {
      double a = 0.0;
      for (int i = 0; i < h; ++i) a += w2[(long)o * h + i] * H1[(long)r * h + i];
      a += b2[o];
      H2[(long)r * h + o] = R(a > 0.0 ? a : 0.0);
    }
    double mx = -INFINITY, se = 0.0;
    for (int o = 0; o < g; ++o) {
      double a = 0.0;
      for (int i = 0; i < h; ++i) a += w3[(long)o * h + i] * H2[(long)r * h + i];
      z[o] = a + b3[o];
      if (z[o] > mx) mx = z[o];
    }
    for (int o = 0; o < g; ++o) se += exp(z[o] - mx);
    loss += log(se) + mx - z[t[r]];
    for (int o = 0; o < g; ++o) {
      const double d = (exp(z[o] - mx) / se - (o == t[r] ? 1.0 : 0.0)) / (double)B;
      dZ[(long)r * g + o] = d;
      dZr[(long)r * g + o] = R(d);
    }
  }
  /* layer 3 */
  for (int o = 0; o < g; ++o)
    for (int i = 0; i < h; ++i) {
      double a = 0.0;
      for (int r = 0; r < B; ++r) a += dZr[(long)r * g + o] * H2[(long)r * h + i];
      gW3[(long)o * h + i] = a;
    }
  for (int o = 0; o < g; ++o)
    for (int r = 0; r < B; ++r) gb3[o] += dZ[(long)r * g + o];
  for (int r = 0; r < B; ++r)
    for (int i = 0; i < h; ++i) {
      double a = 0.0;
      for (int o = 0; o < g; ++o) a += dZr[(long)r * g + o] * w3[(long)o * h + i];
      a = H2[(long)r * h + i] > 0.0 ? a : 0.0;
      dH[(long)r * h + i] = a;
      dHr[(long)r * h + i] = R(a);
    }
  /* layer 2 */
  for (int o = 0; o < h; ++o)
    for (int i = 0; i < h; ++i) {
      double a = 0.0;
      for (int r = 0; r < B; ++r) a += dHr[(long)r * h + o] * H1[(long)r * h + i];
      gW2[(long)o * h + i] = a;
    }
  for (int o = 0; o < h; ++o)
    for (int r = 0; r < B; ++r) gb2[o] += dH[(long)r * h + o];
  double *dH1 = malloc(sizeof(double) * (size_t)B * h), *dH1r = malloc(sizeof(double) * (size_t)B * h);
  for (int r = 0; r < B; ++r)
    for (int i = 0; i < h; ++i) {
      double a = 0.0;
      for (int o = 0; o < h; ++o) a += dHr[(long)r * h + o] * w2[(long)o * h + i];
      a = H1[(long)r * h + i] > 0.0 ? a : 0.0;
      dH1[(long)r * h + i] = a;
      dH1r[(long)r * h + i] = R(a);
    }
  /* layer 1 */
  for (int o = 0; o < h; ++o)
    for (int i = 0; i < din; ++i) {
      double a = 0.0;
      for (int r = 0; r < B; ++r) a += dH1r[(long)r * h + o] * (double)X[(long)r * din + i];
      gW1[(long)o * din + i] = a;
    }
  for (int o = 0; o < h; ++o)
    for (int r = 0; r < B; ++r) gb1[o] += dH1[(long)r * h + o];
#undef R
  free(w1); free(w2); free(w3); free(H1); free(H2); free(dZ); free(dH); free(dHr); free(dZr);
  free(dH1); free(dH1r); free(z);
  return loss / (double)B;
}
