// Drives the kmeans<TYPE> class of include/kmeans_amd.hpp the way a caller of
// the reference's class (main/cpp/kmeans.hpp:174-312) would, for the GPU
// tests: compute() against the step-by-step loop, in float and in double,
// and computeKmeansF called with the reference's default miniBatchFraction.
// usage: class_api_test <points.f32> m n k iters <out.bin>
// out.bin: centres n x k float32 (compute()), labels m int32, compactness
// double, centres n x k float64 (kmeans<double>::compute()), labels m int32
// and compactness double (kmeans<double> step by step), then int32 flags:
// float steps match, float convergence flag, double steps match, double
// convergence flag, computeKmeansF (default last argument) matches compute().
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kmeans_amd.hpp"

template <class T>
struct Run {
  std::vector<T> centers;  // n x k
  std::vector<int> labels;
  double comp = 0.0;
  int conv = 0;
};

// the whole loop through compute()
template <class T>
static int whole(const std::vector<T> &X, const std::vector<T> &C0, int m, int n, int k, int iters,
                 Run<T> *r) {
  kmeans_amd::kmeans<T> a(m, n, k, (T)-1, iters, 1, 1, 0, 0);
  r->centers = C0;
  if (a.buildData(X.data()) || a.initialize(r->centers.data()) || a.compute()) return 1;
  if (a.getClusterCenters(r->centers.data())) return 2;
  r->labels.resize(m);
  if (a.getLabels(r->labels.data())) return 3;
  r->comp = a.getCompactness();
  r->conv = a.getConvergenceFlag();
  return 0;
}

// the same iterations step by step
template <class T>
static int steps(const std::vector<T> &X, const std::vector<T> &C0, int m, int n, int k, int iters,
                 Run<T> *r) {
  kmeans_amd::kmeans<T> b(m, n, k, (T)-1, iters, 1, 1, 0, 0);
  r->centers = C0;
  if (b.buildData(X.data()) || b.initialize(r->centers.data())) return 1;
  for (int it = 0; it < iters; ++it)
    if (b.reset() || b.closestCenters() || b.clusterCenters() || b.compactness() || b.transpose())
      return 2;
  r->comp = b.getCompactness();
  if (b.getClusterCenters(r->centers.data())) return 3;
  r->labels.resize(m);
  if (b.getLabels(r->labels.data())) return 4;
  return 0;
}

template <class T>
static bool same(const Run<T> &a, const Run<T> &b) {
  return a.centers == b.centers && a.labels == b.labels && a.comp == b.comp;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s points.f32 m n k iters out.bin\n", argv[0]);
    return 2;
  }
  const int m = std::atoi(argv[2]), n = std::atoi(argv[3]), k = std::atoi(argv[4]);
  const int iters = std::atoi(argv[5]);
  std::vector<float> X((size_t)m * n);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(X.data(), sizeof(float), X.size(), f) != X.size()) return 3;
  std::fclose(f);
  // initial centres: the first k points, in the reference's n x k layout
  std::vector<float> C0((size_t)n * k);
  for (int j = 0; j < k; ++j)
    for (int q = 0; q < n; ++q) C0[(size_t)q * k + j] = X[(size_t)j * n + q];
  std::vector<double> Xd(X.begin(), X.end()), C0d(C0.begin(), C0.end());

  Run<float> fa, fb;
  Run<double> da, db;
  if (whole(X, C0, m, n, k, iters, &fa)) return 4;
  if (steps(X, C0, m, n, k, iters, &fb)) return 5;
  if (whole(Xd, C0d, m, n, k, iters, &da)) return 6;
  if (steps(Xd, C0d, m, n, k, iters, &db)) return 7;
  const int fmatch = same(fa, fb), dmatch = same(da, db);

  // computeKmeansF as a C++ caller of the reference writes it: the last
  // argument (miniBatchFraction) left to its default of 1
  std::vector<float> rc(C0);
  const kmeansErrorStatus cst =
      computeKmeansF(X.data(), m, n, k, -1.0f, iters, 1, 1, 0, 0, rc.data());
  const int cmatch = (cst == KMEANS_NONCONVERGENCE) && rc == fa.centers;

  FILE *o = std::fopen(argv[6], "wb");
  if (!o) return 13;
  std::fwrite(fa.centers.data(), sizeof(float), fa.centers.size(), o);
  std::fwrite(fa.labels.data(), sizeof(int), fa.labels.size(), o);
  std::fwrite(&fa.comp, sizeof(double), 1, o);
  std::fwrite(da.centers.data(), sizeof(double), da.centers.size(), o);
  std::fwrite(db.labels.data(), sizeof(int), db.labels.size(), o);
  std::fwrite(&db.comp, sizeof(double), 1, o);
  const int flags[5] = {fmatch, fa.conv, dmatch, da.conv, cmatch};
  std::fwrite(flags, sizeof(int), 5, o);
  std::fclose(o);
  std::printf("class api: float steps %s compute(), double steps %s compute()\n",
              fmatch ? "match" : "DIFFER FROM", dmatch ? "match" : "DIFFER FROM");
  return 0;
}
