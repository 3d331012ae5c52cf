// libstdc++'s std::nth_element (bits/stl_algo.h, bits/stl_heap.h: __introselect,
// __unguarded_partition_pivot, __move_median_to_first, __unguarded_partition, __insertion_sort,
// __unguarded_linear_insert, __heap_select, __make_heap, __adjust_heap, __push_heap, __pop_heap)
// restated step by step over an array of kd items ordered by one coordinate, so that the element
// arrangement it leaves - and with it the photon tree KDTree::_Build (Raytracer/Core/KDTree.h:
// 427-473) makes - equals a libstdc++ build. Compiles for the host and the device.
#pragma once
#include <cstddef>
#include <cstdint>

#ifdef __HIPCC__
#define SKW_HD __host__ __device__
#else
#define SKW_HD
#endif

namespace skw {
namespace kdsel {

// A kd item: position and the caller's photon index (16 bytes).
struct alignas(16) Item {
  float p[3];
  uint32_t id;
};

SKW_HD inline float key(const Item &x, int axis) {
  const float v0 = x.p[0], v1 = x.p[1], v2 = x.p[2];
  return axis == 0 ? v0 : (axis == 1 ? v1 : v2);
}

struct Seq {
  Item *a;
  int axis;
  SKW_HD bool less(int64_t i, int64_t j) const { return key(a[i], axis) < key(a[j], axis); }
  SKW_HD bool less_v(const Item &v, int64_t j) const { return key(v, axis) < key(a[j], axis); }
  SKW_HD bool less_iv(int64_t i, const Item &v) const { return key(a[i], axis) < key(v, axis); }
  SKW_HD void swap(int64_t i, int64_t j) const {
    Item t = a[i];
    a[i] = a[j];
    a[j] = t;
  }

  SKW_HD void move_median_to_first(int64_t result, int64_t x, int64_t y, int64_t z) const {
    if (less(x, y)) {
      if (less(y, z)) swap(result, y);
      else if (less(x, z)) swap(result, z);
      else swap(result, x);
    } else if (less(x, z)) {
      swap(result, x);
    } else if (less(y, z)) {
      swap(result, z);
    } else {
      swap(result, y);
    }
  }

  SKW_HD int64_t unguarded_partition(int64_t first, int64_t last, int64_t pivot) const {
    while (true) {
      while (less(first, pivot)) ++first;
      --last;
      while (less(pivot, last)) --last;
      if (!(first < last)) return first;
      swap(first, last);
      ++first;
    }
  }

  SKW_HD int64_t unguarded_partition_pivot(int64_t first, int64_t last) const {
    const int64_t mid = first + (last - first) / 2;
    move_median_to_first(first, first + 1, mid, last - 1);
    return unguarded_partition(first + 1, last, first);
  }

  SKW_HD void unguarded_linear_insert(int64_t last) const {
    const Item val = a[last];
    int64_t next = last - 1;
    while (less_v(val, next)) {
      a[last] = a[next];
      last = next;
      --next;
    }
    a[last] = val;
  }

  SKW_HD void insertion_sort(int64_t first, int64_t last) const {
    if (first == last) return;
    for (int64_t i = first + 1; i != last; ++i) {
      if (less(i, first)) {
        const Item val = a[i];
        for (int64_t k = i; k > first; --k) a[k] = a[k - 1];  // move_backward(first, i, i + 1)
        a[first] = val;
      } else {
        unguarded_linear_insert(i);
      }
    }
  }

  SKW_HD void push_heap(int64_t first, int64_t hole, int64_t top, const Item &value) const {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && less_iv(first + parent, value)) {
      a[first + hole] = a[first + parent];
      hole = parent;
      parent = (hole - 1) / 2;
    }
    a[first + hole] = value;
  }

  SKW_HD void adjust_heap(int64_t first, int64_t hole, int64_t len, const Item &value) const {
    const int64_t top = hole;
    int64_t second = hole;
    while (second < (len - 1) / 2) {
      second = 2 * (second + 1);
      if (less(first + second, first + (second - 1))) second--;
      a[first + hole] = a[first + second];
      hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
      second = 2 * (second + 1);
      a[first + hole] = a[first + (second - 1)];
      hole = second - 1;
    }
    push_heap(first, hole, top, value);
  }

  SKW_HD void make_heap(int64_t first, int64_t last) const {
    if (last - first < 2) return;
    const int64_t len = last - first;
    int64_t parent = (len - 2) / 2;
    while (true) {
      const Item value = a[first + parent];
      adjust_heap(first, parent, len, value);
      if (parent == 0) return;
      parent--;
    }
  }

  SKW_HD void pop_heap(int64_t first, int64_t last, int64_t result) const {
    const Item value = a[result];
    a[result] = a[first];
    adjust_heap(first, 0, last - first, value);
  }

  SKW_HD void heap_select(int64_t first, int64_t middle, int64_t last) const {
    make_heap(first, middle);
    for (int64_t i = middle; i < last; ++i)
      if (less(i, first)) pop_heap(first, middle, i);
  }

  // __introselect from a given depth budget; returns with [first, last) arranged as libstdc++.
  SKW_HD void introselect(int64_t first, int64_t nth, int64_t last, int64_t depth_limit) const {
    while (last - first > 3) {
      if (depth_limit == 0) {
        heap_select(first, nth + 1, last);
        swap(first, nth);
        return;
      }
      --depth_limit;
      const int64_t cut = unguarded_partition_pivot(first, last);
      if (cut <= nth) first = cut;
      else last = cut;
    }
    insertion_sort(first, last);
  }
};

// std::__lg: floor(log2(n)) for n > 0.
SKW_HD inline int64_t lg(uint64_t n) {
  int64_t r = -1;
  while (n) {
    n >>= 1;
    ++r;
  }
  return r;
}

// std::nth_element(a + first, a + nth, a + last) by coordinate `axis`.
SKW_HD inline void nth_element(Item *a, int axis, int64_t first, int64_t nth, int64_t last) {
  if (first == last || nth == last) return;
  Seq s{a, axis};
  s.introselect(first, nth, last, lg((uint64_t)(last - first)) * 2);
}

}  // namespace kdsel
}  // namespace skw
