// Photon kd-tree build on gfx950, identical element for element to KDTree::_Build
// (Raytracer/Core/KDTree.h:427-473) with libstdc++'s std::nth_element (kd_select.hpp).
//
// The tree's shape depends only on n: range [b, e) has the node b, children [b + 1, mid + 1) and
// [mid + 1, e) with mid = (b + e) / 2. The host lists the ranges level by level; the data decide
// the split axis (longest side of the double bbox) and the arrangement nth_element leaves.
// * Large nodes (> 2048 items): one 1024-thread workgroup per node. Each introselect step runs in
//   parallel: after the median of three is moved to the front, the k-th swap of libstdc++'s
//   unguarded partition pairs the k-th position from the left holding a key >= the pivot with the
//   k-th position from the right holding a key <= the pivot, for every k with left < right; the
//   cut is the first left candidate after the last swap, or that swap's right position if it
//   comes first. Candidate positions are counted per thread segment, scanned and written in
//   order; the last swap index is a binary search. Once the subrange fits in LDS it moves there.
// * Nodes of 33-2048 items: one wave per node, the node in LDS, the same parallel steps.
// * Subranges of at most 32 items: one thread finishes them sequentially (and whole subtrees of at
//   most 32 items are built by one thread each, in LDS).
#include <hip/hip_runtime.h>

#include "kd_select.hpp"
#include "kernels.hpp"

namespace skw {
namespace {

using kdsel::Item;
using kdsel::Seq;

constexpr int64_t kLds = (int64_t)kKdLdsItems;
constexpr int64_t kSeq = (int64_t)kKdSeqMax;
constexpr int kBig = 1024;

__device__ __forceinline__ float fminx(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmaxx(float a, float b) { return a < b ? b : a; }

// Exclusive scan of two counts over the G threads of the workgroup (G = 64 or 1024).
template <int G>
__device__ __forceinline__ void scan2(uint32_t a, uint32_t b, uint32_t &ea, uint32_t &eb, uint32_t &ta, uint32_t &tb,
                                      uint32_t *sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t ia = a, ib = b;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t xa = __shfl_up(ia, d), xb = __shfl_up(ib, d);
    if (lane >= d) {
      ia += xa;
      ib += xb;
    }
  }
  if (G == 64) {
    ta = __shfl(ia, 63);
    tb = __shfl(ib, 63);
    ea = ia - a;
    eb = ib - b;
    return;
  }
  constexpr int W = G / 64;
  if (lane == 63) {
    sh[wave] = ia;
    sh[W + wave] = ib;
  }
  __syncthreads();
  uint32_t oa = 0, ob = 0, sa = 0, sb = 0;
  for (int w = 0; w < W; ++w) {
    const uint32_t va = sh[w], vb = sh[W + w];
    if (w < wave) {
      oa += va;
      ob += vb;
    }
    sa += va;
    sb += vb;
  }
  ea = oa + ia - a;
  eb = ob + ib - b;
  ta = sa;
  tb = sb;
  __syncthreads();
}

// One introselect step on [f, l) (l - f > 3) by the G threads of the workgroup: median of three
// to f, then the unguarded partition of [f + 1, l) around it. Positions are kept relative to f in
// sl / sr (Pos wide enough for l - f). Returns the cut (the same value in every thread).
template <int G, class Pos>
__device__ int64_t group_partition(Item *a, Pos *sl, Pos *sr, int axis, int64_t f, int64_t l, uint32_t *sh,
                                   int64_t *sh64) {
  const int t = threadIdx.x;
  if (t == 0) {
    Seq s{a, axis};
    s.move_median_to_first(f, f + 1, f + (l - f) / 2, l - 1);
  }
  __syncthreads();
  // keys of the split axis only: item p's key is float 4 p + axis
  const float *kp = reinterpret_cast<const float *>(a) + axis;
  const float P = kp[4 * f];
  const int64_t m = l - f;
  const int64_t seg = (m + G - 1) / G;
  const int64_t s0 = f + (int64_t)t * seg;
  const int64_t s1 = s0 + seg < l ? s0 + seg : l;
  constexpr int U = G >= 1024 ? 32 : 8;  // keys loaded together (independent loads in flight per thread)
  uint32_t cg = 0, cl = 0;
  for (int64_t p0 = s0; p0 < s1; p0 += U) {
    float k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = p0 + u < s1 ? kp[4 * (p0 + u)] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t p = p0 + u;
      if (p < s1) {
        cg += (p > f && !(k[u] < P)) ? 1u : 0u;
        cl += !(P < k[u]) ? 1u : 0u;
      }
    }
  }
  uint32_t eg, el, ng, nl;
  scan2<G>(cg, cl, eg, el, ng, nl, sh);
  for (int64_t p0 = s0; p0 < s1; p0 += U) {
    float k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = p0 + u < s1 ? kp[4 * (p0 + u)] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t p = p0 + u;
      if (p < s1) {
        if (p > f && !(k[u] < P)) sl[eg++] = (Pos)(p - f);
        if (!(P < k[u])) sr[el++] = (Pos)(p - f);
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    // swaps happen while the k-th left candidate precedes the k-th right one (R[k] = sr[nl - 1 - k])
    uint32_t lo = 0, hi = ng < nl ? ng : nl;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if ((uint32_t)sl[mid] < (uint32_t)sr[nl - 1 - mid]) lo = mid + 1;
      else hi = mid;
    }
    sh64[0] = lo;
  }
  __syncthreads();
  const int64_t K = sh64[0];
  for (int64_t k = t; k < K; k += G) {
    const int64_t i = f + (int64_t)sl[k], j = f + (int64_t)sr[nl - 1 - k];
    const Item x = a[i];
    a[i] = a[j];
    a[j] = x;
  }
  int64_t cut;
  if (K == 0) {
    cut = f + (int64_t)sl[0];
  } else {
    cut = f + (int64_t)sr[nl - K];
    if ((uint32_t)K < ng && f + (int64_t)sl[K] < cut) cut = f + (int64_t)sl[K];
  }
  __syncthreads();
  return cut;
}

// Split axis of a node: the longest side of the double bbox (first of equals: x, y, z).
__device__ __forceinline__ int longest_axis(const float mn[3], const float mx[3]) {
  const double e0 = (double)mx[0] - (double)mn[0], e1 = (double)mx[1] - (double)mn[1],
               e2 = (double)mx[2] - (double)mn[2];
  int axis = 0;
  double ea = e0;
  if (e1 > ea) {
    axis = 1;
    ea = e1;
  }
  if (e2 > ea) axis = 2;
  return axis;
}

__device__ __forceinline__ uint32_t node_word(int axis, uint64_t b, uint64_t e, uint64_t nth) {
  const uint32_t right = nth + 1 < e ? (uint32_t)(nth + 1) : kNoChild;
  return (uint32_t)axis | (b < nth ? 4u : 0u) | (right << 3);
}

// Bbox of a[b, e) reduced over the workgroup (G threads); returns the axis in every thread.
template <int G>
__device__ int group_axis(const Item *a, int64_t b, int64_t e, float *shf, int *sh_axis) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  for (int64_t i = b + t; i < e; i += G) {
    const Item x = a[i];
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminx(mn[c], x.p[c]);
      mx[c] = fmaxx(mx[c], x.p[c]);
    }
  }
  for (int d = 32; d > 0; d >>= 1)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminx(mn[c], __shfl_xor(mn[c], d));
      mx[c] = fmaxx(mx[c], __shfl_xor(mx[c], d));
    }
  constexpr int W = G / 64;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) {
      shf[6 * wave + c] = mn[c];
      shf[6 * wave + 3 + c] = mx[c];
    }
  __syncthreads();
  if (t == 0) {
    float bm[3], bx[3];
    for (int c = 0; c < 3; ++c) {
      bm[c] = shf[c];
      bx[c] = shf[3 + c];
    }
    for (int w = 1; w < W; ++w)
      for (int c = 0; c < 3; ++c) {
        bm[c] = fminx(bm[c], shf[6 * w + c]);
        bx[c] = fmaxx(bx[c], shf[6 * w + 3 + c]);
      }
    *sh_axis = longest_axis(bm, bx);
  }
  __syncthreads();
  return *sh_axis;
}

__global__ void k_kd_init(KdBuildArgs a, const float *pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  Item x;
  x.p[0] = pos[3 * i];
  x.p[1] = pos[3 * i + 1];
  x.p[2] = pos[3 * i + 2];
  x.id = (uint32_t)i;
  static_cast<Item *>(a.items)[i] = x;
}

struct BigShared {
  Item lds[kLds];
  uint16_t lsl[kLds], lsr[kLds];
  uint32_t sh[2 * (kBig / 64)];
  int64_t sh64[1];
  float shf[6 * (kBig / 64)];
  int axis;
};

// The rest of a large node's introselect from state (f, l, depth), then its node word.
__device__ void big_node_finish(const KdBuildArgs &a, BigShared &S, int64_t b, int64_t e, int axis, int64_t f,
                                int64_t l, int64_t depth) {
  Item *items = static_cast<Item *>(a.items);
  const int t = threadIdx.x;
  const int64_t nth = (b + e) / 2;
  Item *lds = S.lds;
  uint16_t *lsl = S.lsl, *lsr = S.lsr;
  uint32_t *sh = S.sh;
  int64_t *sh64 = S.sh64;
  if (l - f > kLds) {
    // depth budget spent on a long subrange: libstdc++'s heap selection, sequentially
    if (t == 0) {
      Seq s{items, axis};
      s.introselect(f, nth, l, depth);
    }
  } else {
    const int64_t m = l - f;
    for (int64_t i = t; i < m; i += kBig) lds[i] = items[f + i];
    __syncthreads();
    int64_t lf = 0, ll = m;
    const int64_t ln = nth - f;
    while (ll - lf > kSeq && depth > 0) {
      --depth;
      const int64_t cut = group_partition<kBig, uint16_t>(lds, lsl, lsr, axis, lf, ll, sh, sh64);
      if (cut <= ln) lf = cut;
      else ll = cut;
    }
    if (t == 0) {
      Seq s{lds, axis};
      s.introselect(lf, ln, ll, depth);
    }
    __syncthreads();
    for (int64_t i = t; i < m; i += kBig) items[f + i] = lds[i];
  }
  __syncthreads();
  if (t == 0) {
    const Item x = items[b];
    items[b] = items[nth];
    items[nth] = x;
    a.word[b] = node_word(axis, (uint64_t)b, (uint64_t)e, (uint64_t)nth);
  }
}

__global__ __launch_bounds__(kBig) void k_kd_big(KdBuildArgs a, const uint64_t *ranges) {
  __shared__ BigShared S;
  Item *items = static_cast<Item *>(a.items);
  const int64_t b = (int64_t)ranges[2 * blockIdx.x], e = (int64_t)ranges[2 * blockIdx.x + 1];
  const int axis = group_axis<kBig>(items, b, e, S.shf, &S.axis);
  const int64_t nth = (b + e) / 2;
  int64_t f = b, l = e, depth = 2 * kdsel::lg((uint64_t)(e - b));
  while (l - f > kLds && depth > 0) {
    --depth;
    const int64_t cut =
        group_partition<kBig, uint32_t>(items, a.scr_l + f, a.scr_r + f, axis, f, l, S.sh, S.sh64);
    if (cut <= nth) f = cut;
    else l = cut;
  }
  big_node_finish(a, S, b, e, axis, f, l, depth);
}

// ---------------------------------------------------------------- many workgroups per node
// A level's large nodes are cut into chunks of kKdChunk positions (chunk_node[j] = node of chunk
// j); per chunk, work holds 8 floats (the bbox part, then counts / offsets as uint32).

__device__ __forceinline__ void chunk_span(const KdStepNode &nd, uint32_t j, int64_t &c0, int64_t &c1) {
  c0 = nd.b + (int64_t)(j - nd.chunk0) * (int64_t)kKdChunk;
  c1 = c0 + (int64_t)kKdChunk < nd.e ? c0 + (int64_t)kKdChunk : nd.e;
}

__global__ __launch_bounds__(kBig) void k_st_bbox(KdBuildArgs a, const KdStepNode *nodes, const uint32_t *chunk_node,
                                                  float *work) {
  __shared__ float shf[6 * (kBig / 64)];
  const uint32_t j = blockIdx.x;
  const KdStepNode nd = nodes[chunk_node[j]];
  int64_t c0, c1;
  chunk_span(nd, j, c0, c1);
  const Item *items = static_cast<const Item *>(a.items);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
  float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
  for (int64_t i = c0 + t; i < c1; i += kBig) {
    const Item x = items[i];
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminx(mn[c], x.p[c]);
      mx[c] = fmaxx(mx[c], x.p[c]);
    }
  }
  for (int d = 32; d > 0; d >>= 1)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminx(mn[c], __shfl_xor(mn[c], d));
      mx[c] = fmaxx(mx[c], __shfl_xor(mx[c], d));
    }
  if (lane == 0)
    for (int c = 0; c < 3; ++c) {
      shf[6 * wave + c] = mn[c];
      shf[6 * wave + 3 + c] = mx[c];
    }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kBig / 64; ++w)
      for (int c = 0; c < 3; ++c) {
        shf[c] = fminx(shf[c], shf[6 * w + c]);
        shf[3 + c] = fmaxx(shf[3 + c], shf[6 * w + 3 + c]);
      }
    for (int c = 0; c < 6; ++c) work[8 * (uint64_t)j + c] = shf[c];
  }
}

// One thread per node: the node's bbox from its chunks' parts, the split axis, the initial state.
__global__ void k_st_axis(KdStepNode *nodes, uint32_t n_nodes, const float *work) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  KdStepNode nd = nodes[i];
  float mn[3], mx[3];
  for (int c = 0; c < 3; ++c) {
    mn[c] = work[8 * (uint64_t)nd.chunk0 + c];
    mx[c] = work[8 * (uint64_t)nd.chunk0 + 3 + c];
  }
  for (uint32_t j = nd.chunk0 + 1; j < nd.chunk0 + nd.nchunks; ++j)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminx(mn[c], work[8 * (uint64_t)j + c]);
      mx[c] = fmaxx(mx[c], work[8 * (uint64_t)j + 3 + c]);
    }
  nd.axis = longest_axis(mn, mx);
  nd.f = nd.b;
  nd.l = nd.e;
  nd.nth = (nd.b + nd.e) / 2;
  nd.depth = 2 * kdsel::lg((uint64_t)(nd.e - nd.b));
  nd.active = 1;
  nodes[i] = nd;
}

// Begin a step: nodes whose subrange fits in LDS (or whose depth budget is spent) leave the loop;
// the others take the median of three to f and record the pivot.
__global__ void k_st_begin(KdBuildArgs a, KdStepNode *nodes, uint32_t n_nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  KdStepNode nd = nodes[i];
  if (!nd.active) return;
  if (nd.l - nd.f <= kLds || nd.depth == 0) {
    nd.active = 0;
  } else {
    --nd.depth;
    Seq s{static_cast<Item *>(a.items), nd.axis};
    s.move_median_to_first(nd.f, nd.f + 1, nd.f + (nd.l - nd.f) / 2, nd.l - 1);
    nd.pivot = kdsel::key(static_cast<const Item *>(a.items)[nd.f], nd.axis);
  }
  nodes[i] = nd;
}

// Per chunk: this thread's contiguous segment of the chunk within [f, l) (16 positions).
__device__ __forceinline__ void st_segment(const KdStepNode &nd, uint32_t j, int64_t &s0, int64_t &s1) {
  int64_t c0, c1;
  chunk_span(nd, j, c0, c1);
  if (c0 < nd.f) c0 = nd.f;
  if (c1 > nd.l) c1 = nd.l;
  constexpr int64_t seg = (int64_t)kKdChunk / kBig;
  s0 = c0 + (int64_t)threadIdx.x * seg;
  s1 = s0 + seg < c1 ? s0 + seg : c1;
  if (s0 > c1) s0 = c1;
}

__global__ __launch_bounds__(kBig) void k_st_count(KdBuildArgs a, const KdStepNode *nodes, const uint32_t *chunk_node,
                                                   uint32_t *work) {
  __shared__ uint32_t sh[2 * (kBig / 64)];
  const uint32_t j = blockIdx.x;
  const KdStepNode nd = nodes[chunk_node[j]];
  if (!nd.active) return;
  int64_t s0, s1;
  st_segment(nd, j, s0, s1);
  const float *kp = reinterpret_cast<const float *>(a.items) + nd.axis;
  const float P = nd.pivot;
  uint32_t cg = 0, cl = 0;
  for (int64_t p = s0; p < s1; ++p) {
    const float k = kp[4 * p];
    cg += (p > nd.f && !(k < P)) ? 1u : 0u;
    cl += !(P < k) ? 1u : 0u;
  }
  uint32_t eg, el, tg, tl;
  scan2<kBig>(cg, cl, eg, el, tg, tl, sh);
  if (threadIdx.x == 0) {
    work[8 * (uint64_t)j] = tg;
    work[8 * (uint64_t)j + 1] = tl;
  }
}

// One workgroup per node: exclusive offsets of the node's chunks, the candidate totals.
__global__ __launch_bounds__(kBig) void k_st_scan(KdStepNode *nodes, uint32_t *work) {
  __shared__ uint32_t sh[2 * (kBig / 64)];
  __shared__ uint32_t carry[2];
  KdStepNode &nd = nodes[blockIdx.x];
  if (!nd.active) return;
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nd.nchunks; base += kBig) {
    const uint32_t j = nd.chunk0 + base + threadIdx.x;
    const bool in = base + threadIdx.x < nd.nchunks;
    const uint32_t cg = in ? work[8 * (uint64_t)j] : 0u, cl = in ? work[8 * (uint64_t)j + 1] : 0u;
    uint32_t eg, el, tg, tl;
    scan2<kBig>(cg, cl, eg, el, tg, tl, sh);
    if (in) {
      work[8 * (uint64_t)j + 2] = carry[0] + eg;
      work[8 * (uint64_t)j + 3] = carry[1] + el;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += tg;
      carry[1] += tl;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nd.nl = carry[0];
    nd.nr = carry[1];
  }
}

__global__ __launch_bounds__(kBig) void k_st_scatter(KdBuildArgs a, const KdStepNode *nodes,
                                                     const uint32_t *chunk_node, const uint32_t *work) {
  __shared__ uint32_t sh[2 * (kBig / 64)];
  const uint32_t j = blockIdx.x;
  const KdStepNode nd = nodes[chunk_node[j]];
  if (!nd.active) return;
  int64_t s0, s1;
  st_segment(nd, j, s0, s1);
  const float *kp = reinterpret_cast<const float *>(a.items) + nd.axis;
  const float P = nd.pivot;
  uint32_t cg = 0, cl = 0;
  for (int64_t p = s0; p < s1; ++p) {
    const float k = kp[4 * p];
    cg += (p > nd.f && !(k < P)) ? 1u : 0u;
    cl += !(P < k) ? 1u : 0u;
  }
  uint32_t eg, el, tg, tl;
  scan2<kBig>(cg, cl, eg, el, tg, tl, sh);
  eg += work[8 * (uint64_t)j + 2];
  el += work[8 * (uint64_t)j + 3];
  uint32_t *sl = a.scr_l + nd.f, *sr = a.scr_r + nd.f;
  for (int64_t p = s0; p < s1; ++p) {
    const float k = kp[4 * p];
    if (p > nd.f && !(k < P)) sl[eg++] = (uint32_t)(p - nd.f);
    if (!(P < k)) sr[el++] = (uint32_t)(p - nd.f);
  }
}

// One thread per node: the swap count (binary search), the cut, the next subrange.
__global__ void k_st_cut(KdBuildArgs a, KdStepNode *nodes, uint32_t n_nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  KdStepNode nd = nodes[i];
  if (!nd.active) return;
  const uint32_t *sl = a.scr_l + nd.f, *sr = a.scr_r + nd.f;
  uint32_t lo = 0, hi = nd.nl < nd.nr ? nd.nl : nd.nr;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (sl[mid] < sr[nd.nr - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  const int64_t K = lo;
  int64_t cut;
  if (K == 0) {
    cut = nd.f + (int64_t)sl[0];
  } else {
    cut = nd.f + (int64_t)sr[nd.nr - K];
    if ((uint32_t)K < nd.nl && nd.f + (int64_t)sl[K] < cut) cut = nd.f + (int64_t)sl[K];
  }
  nd.f_old = nd.f;
  nd.K = K;
  if (cut <= nd.nth) nd.f = cut;
  else nd.l = cut;
  nodes[i] = nd;
}

// Swaps of the last step: chunk j of a node swaps pairs [c kKdChunk, (c + 1) kKdChunk) of its K.
__global__ __launch_bounds__(kBig) void k_st_swap(KdBuildArgs a, const KdStepNode *nodes, const uint32_t *chunk_node) {
  const uint32_t j = blockIdx.x;
  const KdStepNode nd = nodes[chunk_node[j]];
  if (!nd.active) return;
  const int64_t k0 = (int64_t)(j - nd.chunk0) * (int64_t)kKdChunk;
  const int64_t k1 = k0 + (int64_t)kKdChunk < nd.K ? k0 + (int64_t)kKdChunk : nd.K;
  Item *items = static_cast<Item *>(a.items);
  const uint32_t *sl = a.scr_l + nd.f_old, *sr = a.scr_r + nd.f_old;
  for (int64_t k = k0 + threadIdx.x; k < k1; k += kBig) {
    const int64_t p = nd.f_old + (int64_t)sl[k], q = nd.f_old + (int64_t)sr[nd.nr - 1 - k];
    const Item x = items[p];
    items[p] = items[q];
    items[q] = x;
  }
}

__global__ __launch_bounds__(kBig) void k_st_tail(KdBuildArgs a, const KdStepNode *nodes) {
  __shared__ BigShared S;
  const KdStepNode nd = nodes[blockIdx.x];
  big_node_finish(a, S, nd.b, nd.e, nd.axis, nd.f, nd.l, nd.depth);
}

// One wave per node of at most kLds items, the node in LDS (dynamic: items, then two uint16 position
// arrays, for the level's largest node).
__global__ __launch_bounds__(64) void k_kd_mid(KdBuildArgs a, const uint64_t *ranges, uint64_t cap) {
  extern __shared__ __align__(16) unsigned char dyn[];
  Item *lds = reinterpret_cast<Item *>(dyn);
  uint16_t *lsl = reinterpret_cast<uint16_t *>(dyn + cap * sizeof(Item));
  uint16_t *lsr = lsl + cap;
  __shared__ uint32_t sh[2];
  __shared__ int64_t sh64[1];
  __shared__ float shf[6];
  __shared__ int sh_axis;
  Item *items = static_cast<Item *>(a.items);
  const int t = threadIdx.x;
  const int64_t b = (int64_t)ranges[2 * blockIdx.x], e = (int64_t)ranges[2 * blockIdx.x + 1];
  const int64_t m = e - b;
  for (int64_t i = t; i < m; i += 64) lds[i] = items[b + i];
  __syncthreads();
  const int axis = group_axis<64>(lds, 0, m, shf, &sh_axis);
  const int64_t nth = (b + e) / 2 - b;
  int64_t f = 0, l = m, depth = 2 * kdsel::lg((uint64_t)m);
  while (l - f > kSeq && depth > 0) {
    --depth;
    const int64_t cut = group_partition<64, uint16_t>(lds, lsl, lsr, axis, f, l, sh, sh64);
    if (cut <= nth) f = cut;
    else l = cut;
  }
  if (t == 0) {
    Seq s{lds, axis};
    s.introselect(f, nth, l, depth);
    const Item x = lds[0];
    lds[0] = lds[nth];
    lds[nth] = x;
    a.word[b] = node_word(axis, (uint64_t)b, (uint64_t)e, (uint64_t)(b + nth));
  }
  __syncthreads();
  for (int64_t i = t; i < m; i += 64) items[b + i] = lds[i];
}

// One thread per subtree of at most kSeq items, built depth first in the thread's LDS slice.
__global__ __launch_bounds__(64) void k_kd_small(KdBuildArgs a, const uint64_t *ranges, uint32_t nr) {
  __shared__ Item lds[64 * kSeq];
  const uint32_t r = blockIdx.x * 64 + threadIdx.x;
  if (r >= nr) return;
  Item *items = static_cast<Item *>(a.items);
  Item *my = lds + threadIdx.x * kSeq;
  const int64_t b = (int64_t)ranges[2 * r], e = (int64_t)ranges[2 * r + 1];
  const int64_t m = e - b;
  for (int64_t i = 0; i < m; ++i) my[i] = items[b + i];
  int64_t sb[16], se[16];
  int sp = 0;
  sb[sp] = 0;
  se[sp] = m;
  ++sp;
  while (sp > 0) {
    --sp;
    const int64_t x = sb[sp], y = se[sp];
    if (x + 1 == y) {
      a.word[b + x] = 3u;
      continue;
    }
    float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int64_t i = x; i < y; ++i)
      for (int c = 0; c < 3; ++c) {
        mn[c] = fminx(mn[c], my[i].p[c]);
        mx[c] = fmaxx(mx[c], my[i].p[c]);
      }
    const int axis = longest_axis(mn, mx);
    // split index of the absolute range: b + (x + y) / 2 (b is added to both ends)
    const int64_t nth = (x + y) / 2;
    kdsel::nth_element(my, axis, x, nth, y);
    const Item tmp = my[x];
    my[x] = my[nth];
    my[nth] = tmp;
    a.word[b + x] = node_word(axis, (uint64_t)(b + x), (uint64_t)(b + y), (uint64_t)(b + nth));
    if (nth + 1 < y) {
      sb[sp] = nth + 1;
      se[sp] = y;
      ++sp;
    }
    if (x < nth) {
      sb[sp] = x + 1;
      se[sp] = nth + 1;
      ++sp;
    }
  }
  for (int64_t i = 0; i < m; ++i) items[b + i] = my[i];
}

__global__ void k_kd_nodes(KdBuildArgs a, const float *weight, const uint32_t *codes, const uint16_t *dir,
                           const uint16_t *normal, DevKdNode *nodes, DevPhoton *photons, uint32_t *perm,
                           float *psplit) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const Item x = static_cast<const Item *>(a.items)[i];
  DevKdNode nd;
  nd.p[0] = x.p[0];
  nd.p[1] = x.p[1];
  nd.p[2] = x.p[2];
  nd.word = a.word[i];
  nodes[i] = nd;
  perm[i] = x.id;
  DevPhoton ph;
  for (int c = 0; c < 3; ++c) ph.w[c] = weight ? weight[3 * (uint64_t)x.id + c] : 0.0f;
  ph.codes = codes ? codes[x.id]
                    : (uint32_t)(dir ? dir[x.id] : 0) | ((uint32_t)(normal ? normal[x.id] : 0) << 16);
  photons[i] = ph;
  if (i == 0) psplit[0] = 0.0f;
}

// Parent splits: every node but the root has one parent, which writes its entry.
__global__ void k_kd_psplit(KdBuildArgs a, const DevKdNode *nodes, float *psplit) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t w = nodes[i].word, axis = w & 3u;
  if (axis == 3u) return;
  const float sv = axis == 0 ? nodes[i].p[0] : (axis == 1 ? nodes[i].p[1] : nodes[i].p[2]);
  if (w & 4u) psplit[i + 1] = sv;
  if ((uint64_t)(w >> 3) < a.n) psplit[w >> 3] = sv;
}

}  // namespace

hipError_t launch_kd_init(const KdBuildArgs &a, const float *pos, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kd_init, dim3((uint32_t)((a.n + 255) / 256)), dim3(256), 0, stream, a, pos);
  return hipGetLastError();
}

hipError_t launch_kd_big(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, hipStream_t stream) {
  if (n_ranges == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kd_big, dim3(n_ranges), dim3(kBig), 0, stream, a, ranges);
  return hipGetLastError();
}

hipError_t launch_kd_step_init(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, const uint32_t *chunk_node,
                               uint32_t n_chunks, float *work, hipStream_t stream) {
  if (n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_st_bbox, dim3(n_chunks), dim3(kBig), 0, stream, a, nodes, chunk_node, work);
  hipLaunchKernelGGL(k_st_axis, dim3((n_nodes + 63) / 64), dim3(64), 0, stream, nodes, n_nodes, work);
  return hipGetLastError();
}

hipError_t launch_kd_step_begin(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, hipStream_t stream) {
  if (n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_st_begin, dim3((n_nodes + 63) / 64), dim3(64), 0, stream, a, nodes, n_nodes);
  return hipGetLastError();
}

hipError_t launch_kd_step_rest(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, const uint32_t *chunk_node,
                               uint32_t n_chunks, float *work, hipStream_t stream) {
  if (n_nodes == 0) return hipSuccess;
  uint32_t *w = reinterpret_cast<uint32_t *>(work);
  hipLaunchKernelGGL(k_st_count, dim3(n_chunks), dim3(kBig), 0, stream, a, nodes, chunk_node, w);
  hipLaunchKernelGGL(k_st_scan, dim3(n_nodes), dim3(kBig), 0, stream, nodes, w);
  hipLaunchKernelGGL(k_st_scatter, dim3(n_chunks), dim3(kBig), 0, stream, a, nodes, chunk_node, w);
  hipLaunchKernelGGL(k_st_cut, dim3((n_nodes + 63) / 64), dim3(64), 0, stream, a, nodes, n_nodes);
  hipLaunchKernelGGL(k_st_swap, dim3(n_chunks), dim3(kBig), 0, stream, a, nodes, chunk_node);
  return hipGetLastError();
}

hipError_t launch_kd_step_tail(const KdBuildArgs &a, const KdStepNode *nodes, uint32_t n_nodes, hipStream_t stream) {
  if (n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_st_tail, dim3(n_nodes), dim3(kBig), 0, stream, a, nodes);
  return hipGetLastError();
}

hipError_t launch_kd_mid(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, uint64_t max_size,
                         hipStream_t stream) {
  if (n_ranges == 0) return hipSuccess;
  const uint64_t cap = (max_size + 7) & ~uint64_t(7);
  const size_t bytes = cap * sizeof(Item) + 2 * cap * sizeof(uint16_t);
  hipLaunchKernelGGL(k_kd_mid, dim3(n_ranges), dim3(64), bytes, stream, a, ranges, cap);
  return hipGetLastError();
}

hipError_t launch_kd_small(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, hipStream_t stream) {
  if (n_ranges == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kd_small, dim3((n_ranges + 63) / 64), dim3(64), 0, stream, a, ranges, n_ranges);
  return hipGetLastError();
}

hipError_t launch_kd_finish(const KdBuildArgs &a, const float *weight, const uint32_t *codes, const uint16_t *dir,
                            const uint16_t *normal, DevKdNode *nodes, DevPhoton *photons, uint32_t *perm,
                            float *psplit, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint32_t grid = (uint32_t)((a.n + 255) / 256);
  hipLaunchKernelGGL(k_kd_nodes, dim3(grid), dim3(256), 0, stream, a, weight, codes, dir, normal, nodes, photons,
                     perm, psplit);
  hipLaunchKernelGGL(k_kd_psplit, dim3(grid), dim3(256), 0, stream, a, nodes, psplit);
  return hipGetLastError();
}

}  // namespace skw
