// Test helper: the product's population order (pyhea_amd/csrc/hea_sort.h)
// against the standard library's std::sort with the reference engine's
// comparator (pyhea/cpp/accelerate.cpp:47-49) on the same index arrays.
//   sort_check <trials> <seed>  -> prints "ok <cases> <heapsorts> <unstable>"
// or the first mismatch and exits 1.  <unstable> counts the cases where a
// stable sort would have given a different order (so the check is not
// vacuous); <heapsorts> counts runs that exhausted the depth budget.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

static long g_heapsorts = 0;
#define HEA_SORT_ON_HEAPSORT() (++g_heapsorts)
#include "hea_sort.h"

static long g_cases = 0, g_unstable = 0;

static bool check(const std::vector<double>& key) {
  const int n = (int)key.size();
  std::vector<int> a(n), b(n), c(n);
  std::iota(a.begin(), a.end(), 0);
  std::iota(b.begin(), b.end(), 0);
  std::iota(c.begin(), c.end(), 0);
  std::sort(a.begin(), a.end(), [&](int x, int y) { return key[x] < key[y]; });
  hea_sort_keys K{key.data(), nullptr, (int64_t)n};
  hea_sort_like_std(b.data(), n, K);
  std::stable_sort(c.begin(), c.end(), [&](int x, int y) { return key[x] < key[y]; });
  ++g_cases;
  if (a != c) ++g_unstable;
  if (a != b) {
    printf("mismatch n=%d\n", n);
    for (int i = 0; i < n && i < 64; ++i) printf("%d:%d/%d ", i, a[i], b[i]);
    printf("\n");
    return false;
  }
  return true;
}

// McIlroy's adversary ("A killer adversary for quicksort", Software:
// Practice and Experience 29(4), 1999): keys decided lazily during a run of
// std::sort so that its pivots are poor; replaying the keys drives the
// introsort loop into its heapsort fallback.
static std::vector<double> killer(int n) {
  std::vector<int> val(n, n);  // n = "gas" (undecided)
  int solid = 0, candidate = -1;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int x, int y) {
    if (val[x] == n && val[y] == n) {
      if (x == candidate) val[x] = solid++;
      else val[y] = solid++;
    }
    if (val[x] == n) candidate = x;
    else if (val[y] == n) candidate = y;
    return val[x] < val[y];
  });
  std::vector<double> key(n);
  for (int i = 0; i < n; ++i) key[i] = (double)val[i];
  return key;
}

int main(int argc, char** argv) {
  const int trials = argc > 1 ? atoi(argv[1]) : 2000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], 0, 0) : 1);
  for (int t = 0; t < trials; ++t) {
    const int n = t < 400 ? t : (int)(rng() % (t % 10 == 0 ? 5000 : 600));
    const int distinct = 1 + (int)(rng() % (1 + n / (1 + (int)(rng() % 8))));
    std::vector<double> key(n);
    const int shape = (int)(rng() % 4);
    for (int i = 0; i < n; ++i) {
      const double r = (double)(rng() % (uint64_t)distinct);
      key[i] = shape == 0 ? r : shape == 1 ? (double)((i * distinct) / (n ? n : 1))
             : shape == 2 ? (double)(distinct - (i * distinct) / (n ? n : 1)) : (double)((i % 7 == 0) ? r : 1.0);
    }
    if (!check(key)) return 1;
  }
  for (int n : {17, 33, 64, 100, 257, 1000, 4096}) {
    std::vector<double> k = killer(n);
    if (!check(k)) return 1;
    for (double& x : k) x = (double)((long)x / 3);  // the same shape with equal keys
    if (!check(k)) return 1;
  }
  printf("ok %ld %ld %ld\n", g_cases, g_heapsorts, g_unstable);
  return 0;
}
