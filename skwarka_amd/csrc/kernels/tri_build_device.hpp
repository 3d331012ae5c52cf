// Device pieces of the triangle-tree build shared by the per-node kernels (tri_build.hip) and the
// multi-workgroup kernels for the largest nodes (tri_build_wide.hip): the host builder's
// arithmetic (host/builders.cpp, after TriangleAccelerator.cpp:505-617), bin sets, node boxes and
// the children records.
// Every translation unit is compiled with -ffp-contract=off; f32 divisions are correctly rounded.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace skw {
namespace tb {

constexpr uint32_t kMaxTries = 200;  // TriangleAccelerator.h:131
constexpr uint32_t kMaxLeaf = 4;     // TriangleAccelerator.h:128
constexpr uint32_t kMaxDepth = 200;  // TriangleAccelerator.h:134
constexpr uint32_t kBlock = 256;
constexpr uint32_t kSeg = 256;       // segment-tree leaves (>= kMaxTries)
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ float fmin_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin_std(double a, double b) { return (b < a) ? b : a; }

// Float order as unsigned order (no NaN here).
__device__ __forceinline__ uint32_t ord32(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord32(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
constexpr uint32_t kOrdPosInf = 0xFF800000u;  // ord32(+inf)
constexpr uint32_t kOrdNegInf = 0x007FFFFFu;  // ord32(-inf)

// BoxF::area (float) and BoxD::area (double) of host/builders.cpp.
__device__ __forceinline__ float area_f(const float mn[3], const float mx[3]) {
  const float x = fabsf(mx[0] - mn[0]), y = fabsf(mx[1] - mn[1]), z = fabsf(mx[2] - mn[2]);
  return 2.0f * (x * y + x * z + y * z);
}
__device__ __forceinline__ double area_d(const double mn[3], const double mx[3]) {
  const double x = fabs(mx[0] - mn[0]), y = fabs(mx[1] - mn[1]), z = fabs(mx[2] - mn[2]);
  return 2.0 * (x * y + x * z + y * z);
}

__device__ __forceinline__ uint32_t bin_index(double v, uint32_t num_tries) {
  if (!(v < 18446744073709551616.0)) return num_tries;
  const uint64_t b = (uint64_t)v;
  return b < num_tries ? (uint32_t)b : num_tries;
}

// Bins of a triangle box on `axis`: lefts_begin from its max, rights_end from its min.
__device__ __forceinline__ void bins_of(const float *bx, int axis, float lo, double coef, uint32_t T, uint32_t &lb,
                                        uint32_t &re) {
  lb = bin_index((double)(bx[3 + axis] - lo) * coef + 1.0, T);
  re = bin_index((double)(bx[axis] - lo) * coef + 1.0, T);
}

// The candidate count and bin coefficient of a node on one axis.
__device__ __forceinline__ uint32_t num_tries(uint32_t n) { return kMaxTries < 2 * n ? kMaxTries : 2 * n; }
__device__ __forceinline__ double bin_coef(uint32_t T, float mn, float mx) { return (double)((float)T / (mx - mn)); }

// The split position of candidate i.
__device__ __forceinline__ double split_pos(uint32_t i, uint32_t T, float mn, float mx) {
  const float ext = mx - mn;
  const float step = ((float)i * ext) / (float)T;
  return (double)(mn + step);
}

// The host's axis choice: the first least cost in x, y, z order.
__device__ __forceinline__ uint32_t pick_axis(double c0, double c1, double c2) {
  return (c0 <= c1 && c0 <= c2) ? 0u : ((c1 <= c0 && c1 <= c2) ? 1u : 2u);
}

// Partition class of a triangle box: 0 left (below the split), 2 right (above), 1 middle.
__device__ __forceinline__ uint32_t class_of(const float *bx, uint32_t axis, double split) {
  if ((double)bx[3 + axis] < split) return 0u;
  if ((double)bx[axis] > split) return 2u;
  return 1u;
}

__device__ __forceinline__ uint32_t info_depth(uint32_t info) { return info >> 16; }
__device__ __forceinline__ uint32_t info_smask(uint32_t info) { return (info >> 5) & 7u; }
__device__ __forceinline__ bool leaf_by_rule(uint32_t n, uint32_t smask, uint32_t depth) {
  return n <= kMaxLeaf || smask == 7u || depth >= kMaxDepth;
}

// ------------------------------------------------------------------ node box keys

// First minimum / maximum in array order: 64-bit keys (ordered value with -0 as +0, position),
// the least min key and greatest max key name the element that holds each extreme.
struct BoxKeys {
  unsigned long long k[6];
  __device__ void clear() {
    for (int c = 0; c < 3; ++c) {
      k[c] = ~0ull;
      k[3 + c] = 0ull;
    }
  }
  __device__ void add(const float *bx, uint32_t p) {
    for (int c = 0; c < 3; ++c) {
      const float vmn = bx[c] == 0.0f ? 0.0f : bx[c];
      const float vmx = bx[3 + c] == 0.0f ? 0.0f : bx[3 + c];
      const unsigned long long kmn = ((unsigned long long)ord32(vmn) << 32) | p;
      const unsigned long long kmx = ((unsigned long long)ord32(vmx) << 32) | (0xFFFFFFFFu - p);
      k[c] = kmn < k[c] ? kmn : k[c];
      k[3 + c] = kmx > k[3 + c] ? kmx : k[3 + c];
    }
  }
};

__device__ __forceinline__ void keys_merge(unsigned long long *dst, const BoxKeys &k) {
  for (int c = 0; c < 3; ++c) {
    atomicMin(&dst[c], k.k[c]);
    atomicMax(&dst[3 + c], k.k[3 + c]);
  }
}

__device__ __forceinline__ void keys_decode(const unsigned long long *key, const float *box, float mn[3],
                                            float mx[3]) {
  for (int c = 0; c < 3; ++c) {
    const uint32_t pmn = (uint32_t)(key[c] & 0xFFFFFFFFull);
    const uint32_t pmx = 0xFFFFFFFFu - (uint32_t)(key[3 + c] & 0xFFFFFFFFull);
    mn[c] = box[6 * (size_t)pmn + c];
    mx[c] = box[6 * (size_t)pmx + 3 + c];
  }
}

// ------------------------------------------------------------------ bin sets of one axis

// Left bins (by lefts_begin), right bins (by rights_end - 1), straddling ranges as a difference
// array of counts and a segment tree of boxes over bins; boxes as ordered-value unions.
struct Bins {
  uint32_t lbmin[3][kMaxTries], lbmax[3][kMaxTries], rbmin[3][kMaxTries], rbmax[3][kMaxTries];
  uint32_t smin[3][2 * kSeg], smax[3][2 * kSeg];
  uint32_t lcnt[kMaxTries], rcnt[kMaxTries];
  int32_t mdiff[kMaxTries + 1];
};

__device__ __forceinline__ void bins_clear(Bins &B, uint32_t tid, uint32_t nthreads) {
  for (uint32_t i = tid; i < kMaxTries; i += nthreads) {
    for (int c = 0; c < 3; ++c) {
      B.lbmin[c][i] = B.rbmin[c][i] = kOrdPosInf;
      B.lbmax[c][i] = B.rbmax[c][i] = kOrdNegInf;
    }
    B.lcnt[i] = B.rcnt[i] = 0;
  }
  for (uint32_t i = tid; i < 2 * kSeg; i += nthreads)
    for (int c = 0; c < 3; ++c) {
      B.smin[c][i] = kOrdPosInf;
      B.smax[c][i] = kOrdNegInf;
    }
  for (uint32_t i = tid; i <= kMaxTries; i += nthreads) B.mdiff[i] = 0;
}

__device__ __forceinline__ void seg_update(Bins &B, uint32_t node, const float *bx) {
  for (int c = 0; c < 3; ++c) {
    atomicMin(&B.smin[c][node], ord32(bx[c]));
    atomicMax(&B.smax[c][node], ord32(bx[3 + c]));
  }
}

__device__ __forceinline__ void bins_add(Bins &B, const float *bx, uint32_t lbg, uint32_t re, uint32_t T) {
  if (re > 0) {
    for (int c = 0; c < 3; ++c) {
      atomicMin(&B.rbmin[c][re - 1], ord32(bx[c]));
      atomicMax(&B.rbmax[c][re - 1], ord32(bx[3 + c]));
    }
    atomicAdd(&B.rcnt[re - 1], 1u);
  }
  if (lbg < T) {
    for (int c = 0; c < 3; ++c) {
      atomicMin(&B.lbmin[c][lbg], ord32(bx[c]));
      atomicMax(&B.lbmax[c][lbg], ord32(bx[3 + c]));
    }
    atomicAdd(&B.lcnt[lbg], 1u);
  }
  if (re < lbg) {
    atomicAdd(&B.mdiff[re], 1);
    atomicAdd(&B.mdiff[lbg], -1);
    // the boxes of bins [re, lbg): canonical segments of the segment tree
    uint32_t l = re + kSeg, r = lbg + kSeg;
    while (l < r) {
      if (l & 1u) seg_update(B, l++, bx);
      if (r & 1u) seg_update(B, --r, bx);
      l >>= 1;
      r >>= 1;
    }
  }
}

// Adds the entries of one bin set into another (device memory), skipping untouched entries.
__device__ __forceinline__ void bins_merge(Bins &dst, const Bins &src, uint32_t tid, uint32_t nthreads) {
  for (uint32_t i = tid; i < kMaxTries; i += nthreads) {
    if (src.lcnt[i]) {
      atomicAdd(&dst.lcnt[i], src.lcnt[i]);
      for (int c = 0; c < 3; ++c) {
        atomicMin(&dst.lbmin[c][i], src.lbmin[c][i]);
        atomicMax(&dst.lbmax[c][i], src.lbmax[c][i]);
      }
    }
    if (src.rcnt[i]) {
      atomicAdd(&dst.rcnt[i], src.rcnt[i]);
      for (int c = 0; c < 3; ++c) {
        atomicMin(&dst.rbmin[c][i], src.rbmin[c][i]);
        atomicMax(&dst.rbmax[c][i], src.rbmax[c][i]);
      }
    }
  }
  for (uint32_t i = tid; i < 2 * kSeg; i += nthreads)
    if (src.smin[0][i] != kOrdPosInf)
      for (int c = 0; c < 3; ++c) {
        atomicMin(&dst.smin[c][i], src.smin[c][i]);
        atomicMax(&dst.smax[c][i], src.smax[c][i]);
      }
  for (uint32_t i = tid; i <= kMaxTries; i += nthreads)
    if (src.mdiff[i]) atomicAdd(&dst.mdiff[i], src.mdiff[i]);
}

// Cost of candidate i (TriangleAccelerator.cpp:580-605 as restated by the host builder):
// cumulative left / right unions and counts, middle = union over the leaf's segment-tree ancestors.
__device__ __forceinline__ double bins_cost(const Bins &B, uint32_t i, uint32_t T, double inv_area) {
  double lmn[3], lmx[3], rmn[3], rmx[3], mmn[3], mmx[3];
  uint64_t ln = 0, rn = 0;
  int64_t mc = 0;
  for (int c = 0; c < 3; ++c) {
    uint32_t a0 = kOrdPosInf, a1 = kOrdNegInf;
    for (uint32_t j = 0; j <= i; ++j) {
      a0 = B.lbmin[c][j] < a0 ? B.lbmin[c][j] : a0;
      a1 = B.lbmax[c][j] > a1 ? B.lbmax[c][j] : a1;
    }
    lmn[c] = (double)unord32(a0);
    lmx[c] = (double)unord32(a1);
    a0 = kOrdPosInf;
    a1 = kOrdNegInf;
    for (uint32_t j = i; j < T; ++j) {
      a0 = B.rbmin[c][j] < a0 ? B.rbmin[c][j] : a0;
      a1 = B.rbmax[c][j] > a1 ? B.rbmax[c][j] : a1;
    }
    rmn[c] = (double)unord32(a0);
    rmx[c] = (double)unord32(a1);
    a0 = kOrdPosInf;
    a1 = kOrdNegInf;
    for (uint32_t s = i + kSeg; s >= 1; s >>= 1) {
      a0 = B.smin[c][s] < a0 ? B.smin[c][s] : a0;
      a1 = B.smax[c][s] > a1 ? B.smax[c][s] : a1;
    }
    mmn[c] = (double)unord32(a0);
    mmx[c] = (double)unord32(a1);
  }
  for (uint32_t j = 0; j <= i; ++j) ln += B.lcnt[j];
  for (uint32_t j = i; j < T; ++j) rn += B.rcnt[j];
  for (uint32_t j = 0; j <= i; ++j) mc += B.mdiff[j];
  const double p1 = dmin_std(1.0, area_d(lmn, lmx) * inv_area);
  const double p2 = dmin_std(1.0, area_d(mmn, mmx) * inv_area);
  const double p3 = dmin_std(1.0, area_d(rmn, rmx) * inv_area);
  double cost = (double)ln * p1 + (double)mc * p2 + (double)rn * p3;
  if (ln == 0 || rn == 0) cost *= 0.8;
  return cost;
}

// The first least cost over costs[0..T) and its split position (0 and +inf when none is finite).
__device__ __forceinline__ void first_min(const double *costs, uint32_t T, float mn, float mx, double &best,
                                          double &pos) {
  best = __builtin_huge_val();
  uint32_t bi = 0;
  bool any = false;
  for (uint32_t i = 0; i < T; ++i)
    if (costs[i] < best) {
      best = costs[i];
      bi = i;
      any = true;
    }
  pos = any ? split_pos(bi, T, mn, mx) : 0.0;
}

// ------------------------------------------------------------------ block scans

// Exclusive scan of one 64-bit value per thread of a 256-thread block; the total in *total.
__device__ __forceinline__ unsigned long long block_scan(unsigned long long v, unsigned long long *wsum,
                                                         unsigned long long *total) {
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  unsigned long long x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  unsigned long long before = 0, all = 0;
  for (uint32_t k = 0; k < kBlock / 64; ++k) {
    if (k < wv) before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// ------------------------------------------------------------------ node records

// Records the children of node `id` (ranges left / middle / right) and lists them by size.
__device__ inline void emit_children(const TriBuildArgs &a, uint32_t id, uint32_t axis, uint32_t tb, uint32_t mid,
                                     uint32_t right, uint32_t te, uint32_t smask, uint32_t depth) {
  const uint32_t rb[3] = {tb, mid, right}, re[3] = {mid, right, te};
  uint32_t cmask = 0, nch = 0;
  for (int c = 0; c < 3; ++c)
    if (re[c] > rb[c]) {
      cmask |= 1u << c;
      ++nch;
    }
  const uint32_t first = atomicAdd(a.node_count, nch);
  TriBuildNode &n = a.nodes[id];
  n.first_child = first;
  n.info = (axis + 1u) | (cmask << 2) | (smask << 5) | (depth << 16);
  uint32_t k = 0;
  for (int c = 0; c < 3; ++c) {
    if (re[c] <= rb[c]) continue;
    const uint32_t cid = first + k++;
    TriBuildNode &ch = a.nodes[cid];
    ch.tb = rb[c];
    ch.te = re[c];
    ch.first_child = kNone;
    const uint32_t cm = c == 1 ? (smask | (1u << axis)) : smask;
    ch.info = (cm << 5) | ((depth + 1u) << 16);
    const uint32_t sz = re[c] - rb[c];
    if (sz <= a.small_max)
      a.next_small[atomicAdd(a.next_small_n, 1u)] = cid;
    else if (sz > a.wide_min)
      a.next_wide[atomicAdd(a.next_wide_n, 1u)] = cid;
    else
      a.next_big[atomicAdd(a.next_big_n, 1u)] = cid;
  }
}

__device__ __forceinline__ void make_leaf(const TriBuildArgs &a, uint32_t id, uint32_t smask, uint32_t depth) {
  TriBuildNode &n = a.nodes[id];
  n.first_child = kNone;
  n.info = (smask << 5) | (depth << 16);
}

}  // namespace tb
}  // namespace skw
