// The reference's initial-centre rows computed the way its code does it, for
// the CPU tests of kmeansReferenceInitRows: the row shuffle of buildData
// (main/cpp/kmeans.cpp:553-556: std::shuffle with std::default_random_engine
// seeded 0) and the pick of initialize (:876-883: srand(42), rand() % m,
// redrawn while std::binary_search over the points so far finds it), here
// with the process-global srand/rand the reference calls.  Prints the caller
// rows indices[points[i]], one per line.
// usage: reference_init m k
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  const int m = std::atoi(argv[1]), k = std::atoi(argv[2]);
  std::vector<int> indices(m);
  for (int i = 0; i < m; ++i) indices[i] = i;
  std::shuffle(indices.begin(), indices.end(), std::default_random_engine(0u));
  std::vector<int> points;
  std::srand(42);
  for (int i = 0; i < k; ++i) {
    int index = std::rand() % m;
    while (std::binary_search(points.begin(), points.end(), index)) index = std::rand() % m;
    points.push_back(index);
  }
  for (int i = 0; i < k; ++i) std::printf("%d\n", indices[points[i]]);
  return 0;
}
