// Drives the kmeans<TYPE> class of include/kmeans_amd.hpp the way a caller of
// the reference's class (main/cpp/kmeans.hpp:174-312) would, for the GPU
// tests: compute() against the step-by-step loop, and the double class.
// usage: class_api_test <points.f32> m n k iters <out.bin>
// out.bin: centres n x k float32 (compute()), labels m int32, compactness
// double, centres n x k float64 (kmeans<double>), step-by-step match int32.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kmeans_amd.hpp"

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

  // 1. the whole loop
  kmeans_amd::kmeans<float> a(m, n, k, -1.0f, iters, 1, 1, 0, 0);
  std::vector<float> ra(C0);
  if (a.buildData(X.data()) || a.initialize(ra.data()) || a.compute()) return 4;
  if (a.getClusterCenters(ra.data())) return 5;
  std::vector<int> la(m);
  if (a.getLabels(la.data())) return 6;
  const double ca = a.getCompactness();
  const int conv = a.getConvergenceFlag();

  // 2. the same iterations step by step
  kmeans_amd::kmeans<float> b(m, n, k, -1.0f, iters, 1, 1, 0, 0);
  std::vector<float> rb(C0);
  if (b.buildData(X.data()) || b.initialize(rb.data())) return 7;
  double cb = 0.0;
  for (int it = 0; it < iters; ++it) {
    if (b.reset() || b.closestCenters() || b.clusterCenters() || b.compactness() || b.transpose())
      return 8;
  }
  cb = b.getCompactness();
  if (b.getClusterCenters(rb.data())) return 9;
  std::vector<int> lb(m);
  if (b.getLabels(lb.data())) return 10;
  const int match = std::memcmp(ra.data(), rb.data(), sizeof(float) * ra.size()) == 0 &&
                    std::memcmp(la.data(), lb.data(), sizeof(int) * la.size()) == 0 && ca == cb;

  // 3. the double class on the same points
  std::vector<double> Xd(X.begin(), X.end()), rd(C0.begin(), C0.end());
  kmeans_amd::kmeans<double> d(m, n, k, -1.0, iters, 1, 1, 0, 0);
  if (d.buildData(Xd.data()) || d.initialize(rd.data()) || d.compute()) return 11;
  if (d.getClusterCenters(rd.data())) return 12;

  FILE *o = std::fopen(argv[6], "wb");
  if (!o) return 13;
  std::fwrite(ra.data(), sizeof(float), ra.size(), o);
  std::fwrite(la.data(), sizeof(int), la.size(), o);
  std::fwrite(&ca, sizeof(double), 1, o);
  std::fwrite(rd.data(), sizeof(double), rd.size(), o);
  std::fwrite(&match, sizeof(int), 1, o);
  std::fwrite(&conv, sizeof(int), 1, o);
  std::fclose(o);
  std::printf("class api: step-by-step %s compute()\n", match ? "matches" : "DIFFERS FROM");
  return 0;
}
