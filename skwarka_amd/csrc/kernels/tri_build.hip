// Device build of the ternary triangle tree (TriangleAccelerator::TriangleAccelerator, Node::Node
// and _DetermineBestSplit; Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:22-147, 505-729),
// node for node the tree of the host builder (host/builders.cpp), level by level:
//
//  * k_tb_level: one workgroup per node larger than kSmallMax triangles. The node box is the fold
//    of its triangles' boxes in array order (min / max keep the first of equal values, which
//    decides the sign of a zero), found with 64-bit (ordered value, position) atomics. The split
//    search bins the triangles per axis exactly as the host (float bin coefficient, double
//    positions, bins 0..num_tries), left / right bins in LDS, straddling ranges as a difference
//    array of counts and a segment tree of boxes; bin boxes are unions (order-free); costs, the
//    first minimal bin and the float split position use the host's arithmetic. The in-place
//    three-way partition (:650-701) is reproduced exactly in parallel: the loop examines the
//    front elements in order and, after each right element, back elements up to the next
//    non-right one, so each element's examination step follows from prefix counts (right elements
//    in front order, non-right ones in back order) and one count of the examined front elements;
//    left and right elements land in examination order from the two ends; the middle block is a
//    queue that every later left element rotates by one, whose final order is the last m entries
//    of the append sequence, resolved by pointer jumping.
//  * k_tb_small: one thread per node of at most kSmallMax triangles: the same box fold, split
//    search (every candidate evaluated directly) and the sequential partition loop.
// Children are allocated as contiguous id blocks; the host lays the tree out depth first as the
// host builder does.
// Every translation unit is compiled with -ffp-contract=off; f32 divisions are correctly rounded.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace skw {
namespace {

constexpr uint32_t kMaxTries = 200;  // TriangleAccelerator.h:131
constexpr uint32_t kMaxLeaf = 4;     // TriangleAccelerator.h:128
constexpr uint32_t kMaxDepth = 200;  // TriangleAccelerator.h:134
constexpr uint32_t kBlock = 256;
constexpr uint32_t kSeg = 256;       // segment-tree leaves (>= kMaxTries)

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

__device__ __forceinline__ uint32_t info_depth(uint32_t info) { return info >> 16; }
__device__ __forceinline__ uint32_t info_smask(uint32_t info) { return (info >> 5) & 7u; }

// Records the children of node `id` (ranges left / middle / right) and queues them by size.
__device__ void emit_children(const TriBuildArgs &a, uint32_t id, uint32_t axis, uint32_t tb, uint32_t mid,
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
    ch.first_child = 0xFFFFFFFFu;
    const uint32_t cm = c == 1 ? (smask | (1u << axis)) : smask;
    ch.info = (cm << 5) | ((depth + 1u) << 16);
    if (re[c] - rb[c] <= a.small_max)
      a.next_small[atomicAdd(a.next_small_n, 1u)] = cid;
    else
      a.next_big[atomicAdd(a.next_big_n, 1u)] = cid;
  }
}

__device__ __forceinline__ void make_leaf(const TriBuildArgs &a, uint32_t id, uint32_t smask, uint32_t depth) {
  TriBuildNode &n = a.nodes[id];
  n.first_child = 0xFFFFFFFFu;
  n.info = (smask << 5) | (depth << 16);
}

// ------------------------------------------------------------------ one thread per small node

__global__ __launch_bounds__(64) void k_tb_small(TriBuildArgs a) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.list_n) return;
  const uint32_t id = a.list[w];
  TriBuildNode &nd = a.nodes[id];
  const uint32_t tb = nd.tb, te = nd.te, n = te - tb;
  const uint32_t smask = info_smask(nd.info), depth = info_depth(nd.info);
  float mn[3] = {__int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000)};
  float mx[3] = {-mn[0], -mn[1], -mn[2]};
  for (uint32_t p = tb; p < te; ++p) {
    const float *bx = a.box + 6 * (size_t)p;
    for (int c = 0; c < 3; ++c) {
      mn[c] = fmin_std(mn[c], bx[c]);
      mx[c] = fmax_std(mx[c], bx[3 + c]);
    }
  }
  for (int c = 0; c < 3; ++c) {
    nd.bmin[c] = mn[c];
    nd.bmax[c] = mx[c];
  }
  if (n <= kMaxLeaf || smask == 7u || depth >= kMaxDepth) {
    make_leaf(a, id, smask, depth);
    return;
  }
  const double inv_area = 1.0 / (double)area_f(mn, mx);
  double axis_cost[3] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val()};
  double axis_pos[3] = {0.0, 0.0, 0.0};
  for (int axis = 0; axis < 3; ++axis) {
    if (smask & (1u << axis)) continue;
    const uint32_t T = kMaxTries < 2 * n ? kMaxTries : 2 * n;
    const double coef = (double)((float)T / (mx[axis] - mn[axis]));
    const float lo = mn[axis];
    double best = __builtin_huge_val();
    for (uint32_t i = 0; i < T; ++i) {
      // candidate i: left = bins 0..i, right = bins i..T-1, middle = straddling bin boundary i
      double lmn[3], lmx[3], mmn[3], mmx[3], rmn[3], rmx[3];
      for (int c = 0; c < 3; ++c) {
        lmn[c] = mmn[c] = rmn[c] = __builtin_huge_val();
        lmx[c] = mmx[c] = rmx[c] = -__builtin_huge_val();
      }
      uint32_t ln = 0, mc = 0, rn = 0;
      for (uint32_t p = tb; p < te; ++p) {
        const float *bx = a.box + 6 * (size_t)p;
        uint32_t lbg, re;
        bins_of(bx, axis, lo, coef, T, lbg, re);
        if (lbg <= i && lbg < T) {
          ++ln;
          for (int c = 0; c < 3; ++c) {
            lmn[c] = dmin_std(lmn[c], (double)bx[c]);
            lmx[c] = lmx[c] < (double)bx[3 + c] ? (double)bx[3 + c] : lmx[c];
          }
        }
        if (re > 0 && re - 1 >= i) {
          ++rn;
          for (int c = 0; c < 3; ++c) {
            rmn[c] = dmin_std(rmn[c], (double)bx[c]);
            rmx[c] = rmx[c] < (double)bx[3 + c] ? (double)bx[3 + c] : rmx[c];
          }
        }
        if (re <= i && i < lbg) {
          ++mc;
          for (int c = 0; c < 3; ++c) {
            mmn[c] = dmin_std(mmn[c], (double)bx[c]);
            mmx[c] = mmx[c] < (double)bx[3 + c] ? (double)bx[3 + c] : mmx[c];
          }
        }
      }
      const double p1 = dmin_std(1.0, area_d(lmn, lmx) * inv_area);
      const double p2 = dmin_std(1.0, area_d(mmn, mmx) * inv_area);
      const double p3 = dmin_std(1.0, area_d(rmn, rmx) * inv_area);
      double cost = (double)ln * p1 + (double)mc * p2 + (double)rn * p3;
      if (ln == 0 || rn == 0) cost *= 0.8;
      if (cost < best) {
        best = cost;
        const float ext = mx[axis] - mn[axis];
        const float step = ((float)i * ext) / (float)T;
        axis_pos[axis] = (double)(mn[axis] + step);
      }
    }
    axis_cost[axis] = best;
  }
  const uint32_t axis = (axis_cost[0] <= axis_cost[1] && axis_cost[0] <= axis_cost[2])
                            ? 0u
                            : ((axis_cost[1] <= axis_cost[0] && axis_cost[1] <= axis_cost[2]) ? 1u : 2u);
  const double split = axis_pos[axis];
  // the sequential three-way partition (TriangleAccelerator.cpp:650-701)
  uint32_t mid = tb, right = te, i = tb;
  auto swap_el = [&](uint32_t x, uint32_t y) {
    for (int c = 0; c < 6; ++c) {
      const float t = a.box[6 * (size_t)x + c];
      a.box[6 * (size_t)x + c] = a.box[6 * (size_t)y + c];
      a.box[6 * (size_t)y + c] = t;
    }
    const uint32_t t = a.ref[x];
    a.ref[x] = a.ref[y];
    a.ref[y] = t;
  };
  while (i < right) {
    const float *bx = a.box + 6 * (size_t)i;
    if ((double)bx[3 + axis] < split) {
      swap_el(mid, i);
      ++mid;
      ++i;
    } else if ((double)bx[axis] > split) {
      --right;
      swap_el(i, right);
    } else {
      ++i;
    }
  }
  emit_children(a, id, axis, tb, mid, right, te, smask, depth);
}

// ------------------------------------------------------------------ one workgroup per larger node

struct Lds {
  uint32_t lbmin[3][kMaxTries], lbmax[3][kMaxTries], rbmin[3][kMaxTries], rbmax[3][kMaxTries];
  uint32_t smin[3][2 * kSeg], smax[3][2 * kSeg];
  uint32_t lcnt[kMaxTries], rcnt[kMaxTries];
  int32_t mdiff[kMaxTries + 1];
  double cost[kMaxTries];
  unsigned long long key[6];
  unsigned long long wsum[kBlock / 64];
  double axis_cost[3], axis_pos[3];
  uint32_t misc[8];
};

// Exclusive block scan of one 64-bit value per thread; returns the prefix, total in *total.
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

__device__ __forceinline__ void seg_update(Lds &L, uint32_t node, const float *bx) {
  for (int c = 0; c < 3; ++c) {
    atomicMin(&L.smin[c][node], ord32(bx[c]));
    atomicMax(&L.smax[c][node], ord32(bx[3 + c]));
  }
}

__global__ __launch_bounds__(kBlock) void k_tb_level(TriBuildArgs a) {
  __shared__ Lds L;
  const uint32_t tid = threadIdx.x;
  const uint32_t id = a.list[blockIdx.x];
  TriBuildNode &nd = a.nodes[id];
  const uint32_t tb = nd.tb, te = nd.te, n = te - tb;
  const uint32_t smask = info_smask(nd.info), depth = info_depth(nd.info);

  // ---- node box: first minimum / maximum in array order
  if (tid < 6) L.key[tid] = tid < 3 ? ~0ull : 0ull;
  __syncthreads();
  {
    unsigned long long k[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    for (uint32_t p = tb + tid; p < te; p += kBlock) {
      const float *bx = a.box + 6 * (size_t)p;
      for (int c = 0; c < 3; ++c) {
        const float vmn = bx[c] == 0.0f ? 0.0f : bx[c];
        const float vmx = bx[3 + c] == 0.0f ? 0.0f : bx[3 + c];
        const unsigned long long kmn = ((unsigned long long)ord32(vmn) << 32) | p;
        const unsigned long long kmx = ((unsigned long long)ord32(vmx) << 32) | (0xFFFFFFFFu - p);
        k[c] = kmn < k[c] ? kmn : k[c];
        k[3 + c] = kmx > k[3 + c] ? kmx : k[3 + c];
      }
    }
    for (int c = 0; c < 3; ++c) {
      atomicMin(&L.key[c], k[c]);
      atomicMax(&L.key[3 + c], k[3 + c]);
    }
  }
  __syncthreads();
  float mn[3], mx[3];
  for (int c = 0; c < 3; ++c) {
    const uint32_t pmn = (uint32_t)(L.key[c] & 0xFFFFFFFFull);
    const uint32_t pmx = 0xFFFFFFFFu - (uint32_t)(L.key[3 + c] & 0xFFFFFFFFull);
    mn[c] = a.box[6 * (size_t)pmn + c];
    mx[c] = a.box[6 * (size_t)pmx + 3 + c];
  }
  if (tid == 0)
    for (int c = 0; c < 3; ++c) {
      nd.bmin[c] = mn[c];
      nd.bmax[c] = mx[c];
    }
  if (n <= kMaxLeaf || smask == 7u || depth >= kMaxDepth) {
    if (tid == 0) make_leaf(a, id, smask, depth);
    return;
  }
  const double inv_area = 1.0 / (double)area_f(mn, mx);
  const uint32_t T = kMaxTries < 2 * n ? kMaxTries : 2 * n;

  // ---- split search per axis
  for (int axis = 0; axis < 3; ++axis) {
    if (smask & (1u << axis)) {
      if (tid == 0) {
        L.axis_cost[axis] = __builtin_huge_val();
        L.axis_pos[axis] = 0.0;
      }
      continue;
    }
    for (uint32_t i = tid; i < kMaxTries; i += kBlock) {
      for (int c = 0; c < 3; ++c) {
        L.lbmin[c][i] = L.rbmin[c][i] = kOrdPosInf;
        L.lbmax[c][i] = L.rbmax[c][i] = kOrdNegInf;
      }
      L.lcnt[i] = L.rcnt[i] = 0;
    }
    for (uint32_t i = tid; i < 2 * kSeg; i += kBlock)
      for (int c = 0; c < 3; ++c) {
        L.smin[c][i] = kOrdPosInf;
        L.smax[c][i] = kOrdNegInf;
      }
    for (uint32_t i = tid; i <= kMaxTries; i += kBlock) L.mdiff[i] = 0;
    __syncthreads();
    const double coef = (double)((float)T / (mx[axis] - mn[axis]));
    const float lo = mn[axis];
    for (uint32_t p = tb + tid; p < te; p += kBlock) {
      const float *bx = a.box + 6 * (size_t)p;
      uint32_t lbg, re;
      bins_of(bx, axis, lo, coef, T, lbg, re);
      if (re > 0) {
        for (int c = 0; c < 3; ++c) {
          atomicMin(&L.rbmin[c][re - 1], ord32(bx[c]));
          atomicMax(&L.rbmax[c][re - 1], ord32(bx[3 + c]));
        }
        atomicAdd(&L.rcnt[re - 1], 1u);
      }
      if (lbg < T) {
        for (int c = 0; c < 3; ++c) {
          atomicMin(&L.lbmin[c][lbg], ord32(bx[c]));
          atomicMax(&L.lbmax[c][lbg], ord32(bx[3 + c]));
        }
        atomicAdd(&L.lcnt[lbg], 1u);
      }
      if (re < lbg) {
        atomicAdd(&L.mdiff[re], 1);
        atomicAdd(&L.mdiff[lbg], -1);
        // the boxes of bins [re, lbg): canonical segments of the segment tree
        uint32_t l = re + kSeg, r = lbg + kSeg;
        while (l < r) {
          if (l & 1u) seg_update(L, l++, bx);
          if (r & 1u) seg_update(L, --r, bx);
          l >>= 1;
          r >>= 1;
        }
      }
    }
    __syncthreads();
    // candidate costs: cumulative left / right unions, middle = union over the leaf's ancestors
    for (uint32_t i = tid; i < T; i += kBlock) {
      double lmn[3], lmx[3], rmn[3], rmx[3], mmn[3], mmx[3];
      uint64_t ln = 0, rn = 0;
      int64_t mc = 0;
      for (int c = 0; c < 3; ++c) {
        uint32_t a0 = kOrdPosInf, a1 = kOrdNegInf;
        for (uint32_t j = 0; j <= i; ++j) {
          a0 = L.lbmin[c][j] < a0 ? L.lbmin[c][j] : a0;
          a1 = L.lbmax[c][j] > a1 ? L.lbmax[c][j] : a1;
        }
        lmn[c] = (double)unord32(a0);
        lmx[c] = (double)unord32(a1);
        a0 = kOrdPosInf;
        a1 = kOrdNegInf;
        for (uint32_t j = i; j < T; ++j) {
          a0 = L.rbmin[c][j] < a0 ? L.rbmin[c][j] : a0;
          a1 = L.rbmax[c][j] > a1 ? L.rbmax[c][j] : a1;
        }
        rmn[c] = (double)unord32(a0);
        rmx[c] = (double)unord32(a1);
        a0 = kOrdPosInf;
        a1 = kOrdNegInf;
        for (uint32_t s = i + kSeg; s >= 1; s >>= 1) {
          a0 = L.smin[c][s] < a0 ? L.smin[c][s] : a0;
          a1 = L.smax[c][s] > a1 ? L.smax[c][s] : a1;
        }
        mmn[c] = (double)unord32(a0);
        mmx[c] = (double)unord32(a1);
      }
      for (uint32_t j = 0; j <= i; ++j) ln += L.lcnt[j];
      for (uint32_t j = i; j < T; ++j) rn += L.rcnt[j];
      for (uint32_t j = 0; j <= i; ++j) mc += L.mdiff[j];
      const double p1 = dmin_std(1.0, area_d(lmn, lmx) * inv_area);
      const double p2 = dmin_std(1.0, area_d(mmn, mmx) * inv_area);
      const double p3 = dmin_std(1.0, area_d(rmn, rmx) * inv_area);
      double cost = (double)ln * p1 + (double)mc * p2 + (double)rn * p3;
      if (ln == 0 || rn == 0) cost *= 0.8;
      L.cost[i] = cost;
    }
    __syncthreads();
    if (tid == 0) {
      double best = __builtin_huge_val();
      uint32_t bi = 0;
      bool any = false;
      for (uint32_t i = 0; i < T; ++i)
        if (L.cost[i] < best) {
          best = L.cost[i];
          bi = i;
          any = true;
        }
      L.axis_cost[axis] = best;
      double pos = 0.0;
      if (any) {
        const float ext = mx[axis] - mn[axis];
        const float step = ((float)bi * ext) / (float)T;
        pos = (double)(mn[axis] + step);
      }
      L.axis_pos[axis] = pos;
    }
    __syncthreads();
  }
  __syncthreads();  // a masked axis's entries are written without a barrier of their own
  const double c0 = L.axis_cost[0], c1 = L.axis_cost[1], c2 = L.axis_cost[2];
  const uint32_t axis = (c0 <= c1 && c0 <= c2) ? 0u : ((c1 <= c0 && c1 <= c2) ? 1u : 2u);
  const double split = L.axis_pos[axis];

  // ---- exact parallel form of the three-way partition (positions relative to tb)
  uint32_t *cls = a.s_cls + tb, *rf = a.s_rf + tb, *nb = a.s_nb + tb, *G = a.s_G + tb, *Rpos = a.s_Rpos + tb;
  uint32_t *order = a.s_order + tb, *elem = a.s_elem + tb, *ptr = a.s_ptr + tb, *dest = a.s_dest + tb;
  auto class_of = [&](uint32_t k) -> uint32_t {
    const float *bx = a.box + 6 * (size_t)(tb + k);
    if ((double)bx[3 + axis] < split) return 0u;  // left
    if ((double)bx[axis] > split) return 2u;      // right
    return 1u;                                    // middle
  };
  // classes; right elements before each front position; non-right ones before each back position
  unsigned long long carry_r = 0, carry_b = 0, tot = 0;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t k = base + tid;
    uint32_t ck = 3u;
    if (k < n) {
      ck = class_of(k);
      cls[k] = ck;
    }
    const unsigned long long pre = block_scan(k < n && ck == 2u ? 1ull : 0ull, L.wsum, &tot);
    if (k < n) rf[k] = (uint32_t)(carry_r + pre);
    carry_r += tot;
  }
  __threadfence_block();
  __syncthreads();
  const uint32_t n_right = (uint32_t)carry_r, n_nonr = n - n_right;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t j = base + tid;  // back order: element n - 1 - j
    const bool nr = j < n && cls[n - 1 - j] != 2u;
    const unsigned long long pre = block_scan(nr ? 1ull : 0ull, L.wsum, &tot);
    if (j < n) nb[j] = (uint32_t)(carry_b + pre);
    carry_b += tot;
  }
  __threadfence_block();
  __syncthreads();
  // G[g - 1] = back elements in the first g groups (a group ends at a non-right element);
  // Rpos[g - 1] = front position of the g-th right element
  for (uint32_t j = tid; j < n; j += kBlock) {
    if (cls[n - 1 - j] != 2u) G[nb[j]] = j + 1;
    if (cls[j] == 2u) Rpos[rf[j]] = j;
  }
  __threadfence_block();
  __syncthreads();
  auto Gf = [&](uint32_t g) -> uint32_t { return g == 0 ? 0u : (g <= n_nonr ? G[g - 1] : n); };
  // front elements examined: k + G(RF(k)) < n (a prefix)
  if (tid == 0) L.misc[0] = 0;
  __syncthreads();
  {
    uint32_t cnt = 0;
    for (uint32_t k = tid; k < n; k += kBlock) cnt += (k + Gf(rf[k]) < n) ? 1u : 0u;
    atomicAdd(&L.misc[0], cnt);
  }
  __syncthreads();
  const uint32_t K = L.misc[0];
  // (the indices below stay inside the node for every class sequence; the bounds checks only keep
  // a wrong result from leaving it)
  for (uint32_t k = tid; k < K; k += kBlock) {
    const uint32_t q = k + Gf(rf[k]);
    if (q < n) order[q] = k;
  }
  for (uint32_t j = tid; j < n - K; j += kBlock) {
    const uint32_t g1 = nb[j];  // group g = nb[j] + 1 follows the g-th right element
    if (g1 >= n_right) continue;
    const uint32_t q = Rpos[g1] + 1 + j;
    if (q < n) order[q] = n - 1 - j;
  }
  __threadfence_block();
  __syncthreads();
  // examination order: left / right destinations, the middle queue's appends
  unsigned long long c_lr = 0, c_m = 0, c_rot = 0;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t q = base + tid;
    const uint32_t e = q < n ? order[q] : 0u;
    const uint32_t ce = q < n && e < n ? cls[e] : 3u;
    const unsigned long long lr =
        block_scan(ce == 0u ? 1ull : (ce == 2u ? (1ull << 32) : 0ull), L.wsum, &tot);
    const unsigned long long lr_tot = tot;
    const unsigned long long mb = block_scan(ce == 1u ? 1ull : 0ull, L.wsum, &tot);
    const unsigned long long m_tot = tot;
    const unsigned long long m_before = c_m + mb;
    const bool rot = ce == 0u && m_before > 0;
    const unsigned long long rb = block_scan(rot ? 1ull : 0ull, L.wsum, &tot);
    const unsigned long long rot_before = c_rot + rb;
    if (q < n && e < n) {
      const unsigned long long lrb = c_lr + lr;
      if (ce == 0u) dest[e] = (uint32_t)(lrb & 0xFFFFFFFFull);
      if (ce == 2u) dest[e] = n - 1u - (uint32_t)(lrb >> 32);
      const uint32_t t = (uint32_t)(m_before + rot_before);
      if (t >= n) {
      } else if (ce == 1u) {
        elem[t] = e;
        ptr[t] = t;
      } else if (rot) {
        ptr[t] = (uint32_t)rot_before;
      }
    }
    c_lr += lr_tot;
    c_m += m_tot;
    c_rot += tot;
  }
  __threadfence_block();
  __syncthreads();
  const uint32_t n_left = (uint32_t)(c_lr & 0xFFFFFFFFull);
  const uint32_t m = (uint32_t)c_m, r_total = (uint32_t)c_rot, T_app = m + r_total;
  // pointer jumping: every append resolves to the middle element it repeats (each append points
  // at an earlier one or at itself, so 33 rounds resolve any chain of 2^32 appends)
  for (int round = 0; round < 33; ++round) {
    if (tid == 0) L.misc[1] = 0;
    __syncthreads();
    uint32_t changed = 0;
    for (uint32_t t = tid; t < T_app && T_app <= n; t += kBlock) {
      const uint32_t p1 = ptr[t];
      const uint32_t p2 = p1 < T_app ? ptr[p1] : p1;
      if (p2 != p1) {
        ptr[t] = p2;
        changed = 1;
      }
    }
    if (changed) atomicOr(&L.misc[1], 1u);
    __threadfence_block();
    __syncthreads();
    if (L.misc[1] == 0) break;
    __syncthreads();
  }
  for (uint32_t t = r_total + tid; t < T_app && T_app <= n; t += kBlock) {
    const uint32_t p = ptr[t];
    const uint32_t e = p < T_app ? elem[p] : n;
    if (e < n) dest[e] = n_left + (t - r_total);
  }
  __threadfence_block();
  __syncthreads();
  // scatter into the second buffer, then back
  for (uint32_t k = tid; k < n; k += kBlock) {
    if (dest[k] >= n) continue;
    const uint32_t d = tb + dest[k];
    for (int c = 0; c < 6; ++c) a.box2[6 * (size_t)d + c] = a.box[6 * (size_t)(tb + k) + c];
    a.ref2[d] = a.ref[tb + k];
  }
  __threadfence_block();
  __syncthreads();
  for (uint32_t k = tid; k < n; k += kBlock) {
    const uint32_t p = tb + k;
    for (int c = 0; c < 6; ++c) a.box[6 * (size_t)p + c] = a.box2[6 * (size_t)p + c];
    a.ref[p] = a.ref2[p];
  }
  if (tid == 0) emit_children(a, id, axis, tb, tb + n_left, te - n_right, te, smask, depth);
}

}  // namespace

hipError_t launch_tri_build_level(const TriBuildArgs &a, hipStream_t stream) {
  if (a.list_n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tb_level, dim3(a.list_n), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_build_small(const TriBuildArgs &a, hipStream_t stream) {
  if (a.list_n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tb_small, dim3((a.list_n + 63) / 64), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace skw
