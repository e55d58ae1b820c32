// Population order of the reference engine on equal fitness keys.
//
// The reference sorts a population by fitness with
//   std::sort(indices, indices + n, [&](int a, int b) { return f[a] < f[b]; })
// (pyhea/cpp/accelerate.cpp:37-64), after local search (:693) and when it
// merges the old and new populations (:305-330).  std::sort is not stable:
// on equal keys the final order is whatever its algorithm leaves, and since a
// lattice's position selects its random stream in the next phase, that order
// feeds every later decision.  For distinct keys any correct sort gives the
// same order (the GPU ranks by (fitness, position)); when two keys are equal
// the order is reproduced here by running the same algorithm on the index
// array.
//
// The algorithm is introsort (D. R. Musser, "Introspective sorting and
// selection algorithms", Software: Practice and Experience 27(8), 1997) in
// the form GNU libstdc++ (GCC 11.4, the reference's toolchain here) runs it:
//   * introsort loop while a range holds more than 16 elements: with the
//     depth budget (2 floor(log2 n)) spent, heapsort the range; else move the
//     median of (first+1, middle, last-1) to first, partition [first+1, last)
//     around it with the unguarded Hoare scheme, recurse on the right part
//     and continue with the left;
//   * a final insertion sort: guarded over the first 16 elements, unguarded
//     over the rest (every element then has a smaller-or-equal one to its
//     left).
// Heapsort is the binary max-heap with hole sifting (sift the hole down to a
// leaf along the larger child, then push the value back up).  The steps are
// kept exactly because each one decides where equal keys end up; the code is
// checked element for element against the library's std::sort on random
// arrays with many equal keys (tests/native/sort_check.cpp).
#pragma once
#include <stdint.h>

#ifndef HEA_SORT_ON_HEAPSORT
#define HEA_SORT_ON_HEAPSORT() ((void)0)  // test hook: the depth budget ran out
#endif

#if defined(__HIPCC__)
#define HEA_SORT_HD __host__ __device__ __forceinline__
#define HEA_SORT_MEMBER __host__ __device__ __forceinline__
#else
#define HEA_SORT_HD static inline
#define HEA_SORT_MEMBER inline
#endif

// v[0..n) are indices into the concatenation a[0..na) ++ b[...] of two key
// arrays (the old and new populations of a merge); "x before y" is
// key(x) < key(y), the reference's comparator.
struct hea_sort_keys {
  const double* a;
  const double* b;
  int64_t na;
  HEA_SORT_MEMBER double key(int i) const { return i < na ? a[i] : b[i - na]; }
  HEA_SORT_MEMBER bool less(int x, int y) const { return key(x) < key(y); }
};

HEA_SORT_HD void hea_sort_swap(int* v, int64_t i, int64_t j) {
  const int t = v[i];
  v[i] = v[j];
  v[j] = t;
}

// Move the median of v[a], v[b], v[c] (by key) to v[r].
HEA_SORT_HD void hea_sort_median_to(int* v, int64_t r, int64_t a, int64_t b, int64_t c,
                                    const hea_sort_keys& K) {
  if (K.less(v[a], v[b])) {
    if (K.less(v[b], v[c])) hea_sort_swap(v, r, b);
    else if (K.less(v[a], v[c])) hea_sort_swap(v, r, c);
    else hea_sort_swap(v, r, a);
  } else if (K.less(v[a], v[c])) {
    hea_sort_swap(v, r, a);
  } else if (K.less(v[b], v[c])) {
    hea_sort_swap(v, r, c);
  } else {
    hea_sort_swap(v, r, b);
  }
}

// Unguarded Hoare partition of [lo, hi) around the pivot at v[pv]; returns
// the first index of the right part.
HEA_SORT_HD int64_t hea_sort_partition(int* v, int64_t lo, int64_t hi, int64_t pv,
                                       const hea_sort_keys& K) {
  for (;;) {
    while (K.less(v[lo], v[pv])) ++lo;
    --hi;
    while (K.less(v[pv], v[hi])) --hi;
    if (!(lo < hi)) return lo;
    hea_sort_swap(v, lo, hi);
    ++lo;
  }
}

// Heap on v[base .. base+len): place `val` at hole `hole` (heap indices),
// sifting it towards `top`.
HEA_SORT_HD void hea_heap_push(int* v, int64_t base, int64_t hole, int64_t top, int val,
                               const hea_sort_keys& K) {
  int64_t parent = (hole - 1) / 2;
  while (hole > top && K.less(v[base + parent], val)) {
    v[base + hole] = v[base + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  v[base + hole] = val;
}

HEA_SORT_HD void hea_heap_adjust(int* v, int64_t base, int64_t hole, int64_t len, int val,
                                 const hea_sort_keys& K) {
  const int64_t top = hole;
  int64_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (K.less(v[base + child], v[base + child - 1])) --child;
    v[base + hole] = v[base + child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    v[base + hole] = v[base + child - 1];
    hole = child - 1;
  }
  hea_heap_push(v, base, hole, top, val, K);
}

// Heapsort of v[lo, hi): build the heap, then pop the maximum to the end.
HEA_SORT_HD void hea_heapsort(int* v, int64_t lo, int64_t hi, const hea_sort_keys& K) {
  const int64_t len = hi - lo;
  if (len >= 2) {
    for (int64_t parent = (len - 2) / 2;; --parent) {
      hea_heap_adjust(v, lo, parent, len, v[lo + parent], K);
      if (parent == 0) break;
    }
  }
  for (int64_t last = hi; last - lo > 1;) {
    --last;
    const int val = v[last];
    v[last] = v[lo];
    hea_heap_adjust(v, lo, 0, last - lo, val, K);
  }
}

// Insertion of v[i] into the sorted run to its left, with no lower bound
// check (some element to the left is not greater).
HEA_SORT_HD void hea_sort_linear_insert(int* v, int64_t i, const hea_sort_keys& K) {
  const int val = v[i];
  int64_t next = i - 1;
  while (K.less(val, v[next])) {
    v[i] = v[next];
    i = next;
    --next;
  }
  v[i] = val;
}

HEA_SORT_HD void hea_sort_insertion(int* v, int64_t lo, int64_t hi, const hea_sort_keys& K) {
  if (lo == hi) return;
  for (int64_t i = lo + 1; i != hi; ++i) {
    if (K.less(v[i], v[lo])) {
      const int val = v[i];
      for (int64_t k = i; k > lo; --k) v[k] = v[k - 1];
      v[lo] = val;
    } else {
      hea_sort_linear_insert(v, i, K);
    }
  }
}

// Sorts v[0..n) by key exactly as the library's std::sort orders it.
HEA_SORT_HD void hea_sort_like_std(int* v, int64_t n, const hea_sort_keys& K) {
  constexpr int64_t kThreshold = 16;
  if (n <= 1) return;
  int lg = 0;
  while (((int64_t)1 << (lg + 1)) <= n) ++lg;
  // introsort loop; the recursion on the right part is an explicit stack of
  // the left parts still to finish (depth <= 2 lg + 1 <= 64)
  struct Frame {
    int64_t lo, hi;
    int depth;
  };
  Frame stack[66];
  int sp = 0;
  int64_t lo = 0, hi = n;
  int depth = 2 * lg;
  for (;;) {
    while (hi - lo > kThreshold) {
      if (depth == 0) {  // this range is finished by heapsort
        HEA_SORT_ON_HEAPSORT();
        hea_heapsort(v, lo, hi, K);
        break;
      }
      --depth;
      const int64_t mid = lo + (hi - lo) / 2;
      hea_sort_median_to(v, lo, lo + 1, mid, hi - 1, K);
      const int64_t cut = hea_sort_partition(v, lo + 1, hi, lo, K);
      stack[sp++] = Frame{lo, cut, depth};  // continue here after [cut, hi)
      lo = cut;
    }
    if (sp == 0) break;
    const Frame f = stack[--sp];
    lo = f.lo;
    hi = f.hi;
    depth = f.depth;
  }
  if (n > kThreshold) {
    hea_sort_insertion(v, 0, kThreshold, K);
    for (int64_t i = kThreshold; i < n; ++i) hea_sort_linear_insert(v, i, K);
  } else {
    hea_sort_insertion(v, 0, n, K);
  }
}
