// Checks the restated libstdc++ selection (skwarka_amd/csrc/kernels/kd_select.hpp) against
// libstdc++ itself: std::nth_element and std::__introselect with small depth budgets (which
// reach __heap_select), on random keys, keys with many ties and sorted / reversed runs. The
// restatement must leave every element in the same place. Prints "ok" and exits 0 on success.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../skwarka_amd/csrc/kernels/kd_select.hpp"

using skw::kdsel::Item;

static bool same(const std::vector<Item> &a, const std::vector<Item> &b) {
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i].id != b[i].id) return false;
  return true;
}

// The partition step computed the way the device does it in parallel: after the median of three
// is moved to `first`, the k-th swap pairs the k-th position (from the left, in [first + 1, last))
// holding a key >= the pivot with the k-th position (from the right, in [first, last)) holding a
// key <= the pivot, for every k with left < right; the cut is the first left candidate after the
// last swap, or that swap's right position if it comes first.
static int64_t pairs_partition(std::vector<Item> &v, int axis, int64_t first, int64_t last) {
  skw::kdsel::Seq s{v.data(), axis};
  const int64_t mid = first + (last - first) / 2;
  s.move_median_to_first(first, first + 1, mid, last - 1);
  const float P = skw::kdsel::key(v[first], axis);
  std::vector<int64_t> L, R;
  for (int64_t p = first + 1; p < last; ++p)
    if (!(skw::kdsel::key(v[p], axis) < P)) L.push_back(p);
  for (int64_t p = last - 1; p >= first; --p)
    if (!(P < skw::kdsel::key(v[p], axis))) R.push_back(p);
  int64_t K = 0;
  while (K < (int64_t)L.size() && K < (int64_t)R.size() && L[K] < R[K]) ++K;
  for (int64_t k = 0; k < K; ++k) s.swap(L[k], R[k]);
  if (K == 0) return L[0];
  int64_t cut = R[K - 1];
  if (K < (int64_t)L.size() && L[K] < cut) cut = L[K];
  return cut;
}

int main() {
  std::mt19937 rng(12345);
  int cases = 0;
  for (int round = 0; round < 4000; ++round) {
    const size_t n = 1 + rng() % (round < 3000 ? 300 : 5000);
    const int kind = round % 5;
    const int axis = (int)(rng() % 3);
    std::vector<Item> v(n);
    for (size_t i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) {
        float x;
        if (kind == 0) x = std::uniform_real_distribution<float>(-1.0f, 1.0f)(rng);
        else if (kind == 1) x = (float)(rng() % 7);                       // heavy ties
        else if (kind == 2) x = (float)i;                                 // sorted
        else if (kind == 3) x = (float)(n - i);                           // reversed
        else x = (float)((i * 7919) % 13) + (rng() % 2 ? 0.5f : 0.0f);   // periodic
        v[i].p[c] = x;
      }
      v[i].id = (uint32_t)i;
    }
    const int64_t nth = (int64_t)(rng() % n);
    auto comp = [axis](const Item &x, const Item &y) { return skw::kdsel::key(x, axis) < skw::kdsel::key(y, axis); };
    // the full call
    std::vector<Item> a = v, b = v;
    std::nth_element(a.begin(), a.begin() + nth, a.end(), comp);
    skw::kdsel::nth_element(b.data(), axis, 0, nth, (int64_t)n);
    if (!same(a, b)) {
      std::printf("nth_element mismatch: round %d n %zu kind %d nth %lld\n", round, n, kind, (long long)nth);
      return 1;
    }
    // small depth budgets: the heap-select fallback
    for (int depth = 0; depth < 4; ++depth) {
      std::vector<Item> c = v, d = v;
      if (n > 0)
        std::__introselect(c.begin(), c.begin() + nth, c.end(), (long)depth,
                           __gnu_cxx::__ops::__iter_comp_iter(comp));
      skw::kdsel::Seq s{d.data(), axis};
      s.introselect(0, nth, (int64_t)n, depth);
      if (!same(c, d)) {
        std::printf("introselect mismatch: round %d n %zu kind %d nth %lld depth %d\n", round, n, kind,
                    (long long)nth, depth);
        return 1;
      }
    }
    // one partition step: the parallel pairing equals the sequential scan
    if (n > 3) {
      std::vector<Item> e = v, f = v;
      skw::kdsel::Seq s{e.data(), axis};
      const int64_t c1 = s.unguarded_partition_pivot(0, (int64_t)n);
      const int64_t c2 = pairs_partition(f, axis, 0, (int64_t)n);
      if (c1 != c2 || !same(e, f)) {
        std::printf("partition mismatch: round %d n %zu kind %d cut %lld vs %lld\n", round, n, kind, (long long)c1,
                    (long long)c2);
        return 1;
      }
    }
    ++cases;
  }
  std::printf("ok %d\n", cases);
  return 0;
}
