// The largest nodes of a triangle-tree level (TriangleAccelerator.cpp:505-729, the steps of
// k_tb_level in tri_build.hip), each spread over chunks of triangles, one workgroup per chunk:
//   box       first-extreme keys per chunk, merged per node with 64-bit atomics;
//   bins      per chunk and axis, LDS bin sets merged into the node's bin sets;
//   split     one workgroup per node: costs, first least cost, axis (the host's choice);
//   partition the exact parallel form of the in-place three-way loop (:650-701): classes and
//             per-chunk counts, chunk scans per node, element prefixes, the examined-front count K
//             (a binary search: the examination step k + G(RF(k)) grows with k), the examination
//             order, its chunk counts and scans, destinations of left / right elements and the
//             middle queue's appends, pointer jumping (in place, a fixed number of rounds: each
//             round at least halves every remaining chain), middle destinations, scatter, copy;
//   emit      the children records, listed by size.
// Every translation unit is compiled with -ffp-contract=off; f32 divisions are correctly rounded.
#include <hip/hip_runtime.h>

#include "tri_build_device.hpp"

namespace skw {
namespace {

using namespace tb;

__device__ __forceinline__ Bins *bins_of_node(const TriWideArgs &a, uint32_t w, uint32_t axis) {
  return reinterpret_cast<Bins *>(a.bins) + 3 * (size_t)w + axis;
}

// The chunk's node and its element range [b, e) relative to the node.
struct ChunkRange {
  uint32_t w, b, e;
};
__device__ __forceinline__ ChunkRange chunk_range(const TriWideArgs &a, uint32_t c) {
  ChunkRange r;
  r.w = a.chunk_w[c];
  const TriWide &W = a.w[r.w];
  r.b = (c - W.c_begin) * a.chunk;
  r.e = r.b + a.chunk < W.n ? r.b + a.chunk : W.n;
  return r;
}

__global__ void k_tbw_setup(TriWideArgs a) {
  if (threadIdx.x != 0) return;
  uint32_t nc = 0, nmax = 0;
  for (uint32_t w = 0; w < a.b.list_n; ++w) {
    const uint32_t id = a.b.list[w];
    const TriBuildNode nd = a.b.nodes[id];
    TriWide &W = a.w[w];
    W.id = id;
    W.tb = nd.tb;
    W.n = nd.te - nd.tb;
    W.info = nd.info;
    const uint32_t k = (W.n + a.chunk - 1) / a.chunk;
    W.c_begin = nc;
    W.c_end = nc + k < a.chunk_cap ? nc + k : a.chunk_cap;
    for (uint32_t c = W.c_begin; c < W.c_end; ++c) a.chunk_w[c] = w;
    nc = W.c_end;
    W.leaf = 1;  // until the split step decides
    nmax = W.n > nmax ? W.n : nmax;
  }
  a.info_out[0] = nc;
  a.info_out[1] = nmax;
}

__global__ __launch_bounds__(kBlock) void k_tbw_init(TriWideArgs a) {
  const uint32_t w = blockIdx.x, axis = blockIdx.y;
  bins_clear(*bins_of_node(a, w, axis), threadIdx.x, kBlock);
  if (axis == 0 && threadIdx.x < 6) a.w[w].key[threadIdx.x] = threadIdx.x < 3 ? ~0ull : 0ull;
}

__global__ __launch_bounds__(kBlock) void k_tbw_box(TriWideArgs a) {
  __shared__ unsigned long long key[6];
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (threadIdx.x < 6) key[threadIdx.x] = threadIdx.x < 3 ? ~0ull : 0ull;
  __syncthreads();
  BoxKeys k;
  k.clear();
  for (uint32_t i = r.b + threadIdx.x; i < r.e; i += kBlock) {
    const uint32_t p = W.tb + i;
    k.add(a.b.box + 6 * (size_t)p, p);
  }
  keys_merge(key, k);
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&a.w[r.w].key[threadIdx.x], key[threadIdx.x]);
  else if (threadIdx.x < 6) atomicMax(&a.w[r.w].key[threadIdx.x], key[threadIdx.x]);
}

__global__ __launch_bounds__(kBlock) void k_tbw_bins(TriWideArgs a) {
  __shared__ Bins B;
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const uint32_t axis = blockIdx.y;
  const TriWide &W = a.w[r.w];
  const uint32_t smask = info_smask(W.info), depth = info_depth(W.info);
  if (leaf_by_rule(W.n, smask, depth) || (smask & (1u << axis))) return;
  float mn[3], mx[3];
  keys_decode(W.key, a.b.box, mn, mx);
  const uint32_t T = num_tries(W.n);
  const double coef = bin_coef(T, mn[axis], mx[axis]);
  const float lo = mn[axis];
  bins_clear(B, threadIdx.x, kBlock);
  __syncthreads();
  for (uint32_t i = r.b + threadIdx.x; i < r.e; i += kBlock) {
    const float *bx = a.b.box + 6 * (size_t)(W.tb + i);
    uint32_t lbg, re;
    bins_of(bx, axis, lo, coef, T, lbg, re);
    bins_add(B, bx, lbg, re, T);
  }
  __syncthreads();
  bins_merge(*bins_of_node(a, r.w, axis), B, threadIdx.x, kBlock);
}

__global__ __launch_bounds__(kBlock) void k_tbw_split(TriWideArgs a) {
  __shared__ Bins B;
  __shared__ double cost[kMaxTries];
  __shared__ double axis_cost[3], axis_pos[3];
  const uint32_t w = blockIdx.x, tid = threadIdx.x;
  TriWide &W = a.w[w];
  const uint32_t smask = info_smask(W.info), depth = info_depth(W.info);
  float mn[3], mx[3];
  keys_decode(W.key, a.b.box, mn, mx);
  TriBuildNode &nd = a.b.nodes[W.id];
  if (tid == 0)
    for (int c = 0; c < 3; ++c) {
      nd.bmin[c] = mn[c];
      nd.bmax[c] = mx[c];
    }
  if (leaf_by_rule(W.n, smask, depth)) {
    if (tid == 0) {
      make_leaf(a.b, W.id, smask, depth);
      W.leaf = 1;
    }
    return;
  }
  const double inv_area = 1.0 / (double)area_f(mn, mx);
  const uint32_t T = num_tries(W.n);
  for (int axis = 0; axis < 3; ++axis) {
    if (smask & (1u << axis)) {
      if (tid == 0) {
        axis_cost[axis] = __builtin_huge_val();
        axis_pos[axis] = 0.0;
      }
      continue;
    }
    {
      const uint32_t *src = reinterpret_cast<const uint32_t *>(bins_of_node(a, w, axis));
      uint32_t *dst = reinterpret_cast<uint32_t *>(&B);
      for (uint32_t i = tid; i < sizeof(Bins) / 4; i += kBlock) dst[i] = src[i];
    }
    __syncthreads();
    for (uint32_t i = tid; i < T; i += kBlock) cost[i] = bins_cost(B, i, T, inv_area);
    __syncthreads();
    if (tid == 0) first_min(cost, T, mn[axis], mx[axis], axis_cost[axis], axis_pos[axis]);
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0) {
    const uint32_t axis = pick_axis(axis_cost[0], axis_cost[1], axis_cost[2]);
    W.axis = axis;
    W.split = axis_pos[axis];
    W.leaf = 0;
  }
}

// classes; per chunk: right elements, non-right elements
__global__ __launch_bounds__(kBlock) void k_tbw_class(TriWideArgs a) {
  __shared__ uint32_t cnt[2];
  const uint32_t c = blockIdx.x;
  const ChunkRange r = chunk_range(a, c);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
  __syncthreads();
  uint32_t nr = 0, nn = 0;
  for (uint32_t i = r.b + threadIdx.x; i < r.e; i += kBlock) {
    const uint32_t cl = class_of(a.b.box + 6 * (size_t)(W.tb + i), W.axis, W.split);
    a.b.s_cls[W.tb + i] = cl;
    nr += cl == 2u ? 1u : 0u;
    nn += cl != 2u ? 1u : 0u;
  }
  atomicAdd(&cnt[0], nr);
  atomicAdd(&cnt[1], nn);
  __syncthreads();
  if (threadIdx.x < 2) a.ccnt[8 * (size_t)c + threadIdx.x] = cnt[threadIdx.x];
}

// Chunk scans of one node: mode 0 after the classes (right elements in front order, non-right
// ones in back order), mode 1 after the examination counts (left, right, middle; then the rotation
// count of each chunk: all its left elements when a middle element comes before the chunk, else
// those after its first middle element).
__global__ __launch_bounds__(kBlock) void k_tbw_scan(TriWideArgs a, int mode) {
  __shared__ unsigned long long wsum[kBlock / 64];
  TriWide &W = a.w[blockIdx.x];
  if (W.leaf) return;
  const uint32_t cb = W.c_begin, ce = W.c_end, nc = ce - cb, tid = threadIdx.x;
  unsigned long long tot = 0;
  if (mode == 0) {
    unsigned long long cr = 0, cn = 0;
    for (uint32_t base = 0; base < nc; base += kBlock) {
      const uint32_t i = base + tid;
      const unsigned long long vr = i < nc ? a.ccnt[8 * (size_t)(cb + i) + 0] : 0ull;
      const unsigned long long pr = block_scan(vr, wsum, &tot);
      if (i < nc) a.cbase[8 * (size_t)(cb + i) + 0] = (uint32_t)(cr + pr);
      cr += tot;
      // back order: chunk ce - 1 - i; base = non-right elements in the chunks after it
      const unsigned long long vn = i < nc ? a.ccnt[8 * (size_t)(ce - 1 - i) + 1] : 0ull;
      const unsigned long long pn = block_scan(vn, wsum, &tot);
      if (i < nc) a.cbase[8 * (size_t)(ce - 1 - i) + 1] = (uint32_t)(cn + pn);
      cn += tot;
    }
    if (tid == 0) {
      W.n_right = (uint32_t)cr;
      W.n_nonr = (uint32_t)cn;
    }
    return;
  }
  unsigned long long cl = 0, cr = 0, cm = 0;
  for (uint32_t base = 0; base < nc; base += kBlock) {
    const uint32_t i = base + tid;
    const size_t o = 8 * (size_t)(cb + i);
    const unsigned long long pl = block_scan(i < nc ? a.ccnt[o + 2] : 0ull, wsum, &tot);
    const unsigned long long tl = tot;
    const unsigned long long pr = block_scan(i < nc ? a.ccnt[o + 3] : 0ull, wsum, &tot);
    const unsigned long long tr = tot;
    const unsigned long long pm = block_scan(i < nc ? a.ccnt[o + 4] : 0ull, wsum, &tot);
    if (i < nc) {
      a.cbase[o + 2] = (uint32_t)(cl + pl);
      a.cbase[o + 3] = (uint32_t)(cr + pr);
      a.cbase[o + 4] = (uint32_t)(cm + pm);
    }
    cl += tl;
    cr += tr;
    cm += tot;
  }
  __syncthreads();
  unsigned long long crot = 0;
  for (uint32_t base = 0; base < nc; base += kBlock) {
    const uint32_t i = base + tid;
    const size_t o = 8 * (size_t)(cb + i);
    unsigned long long v = 0;
    if (i < nc) v = a.cbase[o + 4] > 0 ? a.ccnt[o + 2] : a.ccnt[o + 5];
    const unsigned long long p = block_scan(v, wsum, &tot);
    if (i < nc) a.cbase[o + 5] = (uint32_t)(crot + p);
    crot += tot;
  }
  if (tid == 0) {
    W.n_left = (uint32_t)cl;
    W.m = (uint32_t)cm;
    W.r_total = (uint32_t)crot;
  }
}

// element prefixes: RF(k), NB(j) for back position j = n-1-k, and the G / Rpos tables
__global__ __launch_bounds__(kBlock) void k_tbw_prefix(TriWideArgs a) {
  __shared__ unsigned long long wsum[kBlock / 64];
  const uint32_t c = blockIdx.x;
  const ChunkRange r = chunk_range(a, c);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, tb = W.tb;
  uint32_t *cls = a.b.s_cls + tb, *rf = a.b.s_rf + tb, *nb = a.b.s_nb + tb, *G = a.b.s_G + tb,
           *Rpos = a.b.s_Rpos + tb;
  const uint32_t base_r = a.cbase[8 * (size_t)c + 0];
  const uint32_t after_n = a.cbase[8 * (size_t)c + 1], chunk_n = a.ccnt[8 * (size_t)c + 1];
  unsigned long long carry_r = 0, carry_n = 0, tot = 0;
  for (uint32_t i0 = r.b; i0 < r.e; i0 += kBlock) {
    const uint32_t k = i0 + threadIdx.x;
    const uint32_t cl = k < r.e ? cls[k] : 3u;
    // (R count, non-R count) packed
    const unsigned long long v = cl == 2u ? 1ull : (cl < 2u ? (1ull << 32) : 0ull);
    const unsigned long long p = block_scan(v, wsum, &tot);
    if (k < r.e) {
      const uint32_t rfk = base_r + (uint32_t)(carry_r + (p & 0xFFFFFFFFull));
      rf[k] = rfk;
      const uint32_t nonr_incl = (uint32_t)(carry_n + (p >> 32)) + (cl != 2u ? 1u : 0u);
      const uint32_t j = n - 1 - k;
      const uint32_t nbj = after_n + (chunk_n - nonr_incl);
      nb[j] = nbj;
      if (cl != 2u && nbj < n) G[nbj] = j + 1;
      if (cl == 2u && rfk < n) Rpos[rfk] = k;
    }
    carry_r += tot & 0xFFFFFFFFull;
    carry_n += tot >> 32;
  }
}

__device__ __forceinline__ uint32_t g_of(const uint32_t *G, uint32_t g, uint32_t n_nonr, uint32_t n) {
  return g == 0 ? 0u : (g <= n_nonr ? G[g - 1] : n);
}

// K: front elements examined, the first k with k + G(RF(k)) >= n (the step grows with k)
__global__ void k_tbw_K(TriWideArgs a) {
  if (threadIdx.x != 0) return;
  TriWide &W = a.w[blockIdx.x];
  if (W.leaf) return;
  const uint32_t n = W.n;
  const uint32_t *rf = a.b.s_rf + W.tb, *G = a.b.s_G + W.tb;
  uint32_t lo = 0, hi = n;  // answer in [lo, hi]
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (mid + g_of(G, rf[mid], W.n_nonr, n) >= n)
      hi = mid;
    else
      lo = mid + 1;
  }
  W.K = lo;
}

// the examination order
__global__ __launch_bounds__(kBlock) void k_tbw_order(TriWideArgs a) {
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, tb = W.tb, K = W.K;
  const uint32_t *rf = a.b.s_rf + tb, *nb = a.b.s_nb + tb, *G = a.b.s_G + tb, *Rpos = a.b.s_Rpos + tb;
  uint32_t *order = a.b.s_order + tb;
  // (the indices below stay inside the node for every class sequence; the bounds checks only keep
  // a wrong result from leaving it)
  for (uint32_t k = r.b + threadIdx.x; k < r.e; k += kBlock) {
    if (k < K) {
      const uint32_t q = k + g_of(G, rf[k], W.n_nonr, n);
      if (q < n) order[q] = k;
    } else {
      const uint32_t j = n - 1 - k;
      const uint32_t g1 = nb[j];
      if (g1 >= W.n_right) continue;
      const uint32_t q = Rpos[g1] + 1 + j;
      if (q < n) order[q] = k;
    }
  }
}

// per chunk of the examination order: left, right, middle; left after the chunk's first middle
__global__ __launch_bounds__(kBlock) void k_tbw_xcount(TriWideArgs a) {
  __shared__ uint32_t cnt[4], first_m;
  const uint32_t c = blockIdx.x;
  const ChunkRange r = chunk_range(a, c);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, tb = W.tb;
  const uint32_t *order = a.b.s_order + tb, *cls = a.b.s_cls + tb;
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) first_m = 0xFFFFFFFFu;
  __syncthreads();
  uint32_t l = 0, rr = 0, m = 0, fm = 0xFFFFFFFFu;
  for (uint32_t q = r.b + threadIdx.x; q < r.e; q += kBlock) {
    const uint32_t e = order[q];
    const uint32_t ce = e < n ? cls[e] : 3u;
    l += ce == 0u;
    rr += ce == 2u;
    m += ce == 1u;
    if (ce == 1u && q < fm) fm = q;
  }
  atomicAdd(&cnt[0], l);
  atomicAdd(&cnt[1], rr);
  atomicAdd(&cnt[2], m);
  atomicMin(&first_m, fm);
  __syncthreads();
  uint32_t la = 0;
  for (uint32_t q = r.b + threadIdx.x; q < r.e; q += kBlock) {
    const uint32_t e = order[q];
    if (q > first_m && e < n && cls[e] == 0u) ++la;
  }
  atomicAdd(&cnt[3], la);
  __syncthreads();
  if (threadIdx.x < 4) a.ccnt[8 * (size_t)c + 2 + threadIdx.x] = cnt[threadIdx.x];
}

// destinations of left / right elements, the middle queue's appends
__global__ __launch_bounds__(kBlock) void k_tbw_exam(TriWideArgs a) {
  __shared__ unsigned long long wsum[kBlock / 64];
  const uint32_t c = blockIdx.x;
  const ChunkRange r = chunk_range(a, c);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, tb = W.tb;
  const uint32_t *order = a.b.s_order + tb, *cls = a.b.s_cls + tb;
  uint32_t *dest = a.b.s_dest + tb, *elem = a.b.s_elem + tb, *ptr = a.b.s_ptr + tb;
  const size_t o = 8 * (size_t)c;
  unsigned long long c_l = a.cbase[o + 2], c_r = a.cbase[o + 3], c_m = a.cbase[o + 4], c_rot = a.cbase[o + 5];
  unsigned long long tot = 0;
  for (uint32_t q0 = r.b; q0 < r.e; q0 += kBlock) {
    const uint32_t q = q0 + threadIdx.x;
    const uint32_t e = q < r.e ? order[q] : 0u;
    const uint32_t ce = q < r.e && e < n ? cls[e] : 3u;
    const unsigned long long lr = block_scan(ce == 0u ? 1ull : (ce == 2u ? (1ull << 32) : 0ull), wsum, &tot);
    const unsigned long long lr_tot = tot;
    const unsigned long long mb = block_scan(ce == 1u ? 1ull : 0ull, wsum, &tot);
    const unsigned long long m_tot = tot;
    const unsigned long long m_before = c_m + mb;
    const bool rot = ce == 0u && m_before > 0;
    const unsigned long long rb = block_scan(rot ? 1ull : 0ull, wsum, &tot);
    const unsigned long long rot_before = c_rot + rb;
    if (q < r.e && e < n) {
      if (ce == 0u) dest[e] = (uint32_t)(c_l + (lr & 0xFFFFFFFFull));
      if (ce == 2u) dest[e] = n - 1u - (uint32_t)(c_r + (lr >> 32));
      const uint32_t t = (uint32_t)(m_before + rot_before);
      if (t >= n) {
      } else if (ce == 1u) {
        elem[t] = e;
        ptr[t] = t;
      } else if (rot) {
        ptr[t] = (uint32_t)rot_before;
      }
    }
    c_l += lr_tot & 0xFFFFFFFFull;
    c_r += lr_tot >> 32;
    c_m += m_tot;
    c_rot += tot;
  }
}

// one round of pointer jumping over the appends of every node
__global__ __launch_bounds__(kBlock) void k_tbw_jump(TriWideArgs a) {
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t T_app = W.m + W.r_total;
  if (T_app > W.n) return;
  uint32_t *ptr = a.b.s_ptr + W.tb;
  for (uint32_t t = r.b + threadIdx.x; t < r.e && t < T_app; t += kBlock) {
    const uint32_t p1 = ptr[t];
    const uint32_t p2 = p1 < T_app ? ptr[p1] : p1;
    if (p2 != p1) ptr[t] = p2;
  }
}

// middle destinations, then the scatter
__global__ __launch_bounds__(kBlock) void k_tbw_mdest(TriWideArgs a) {
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, T_app = W.m + W.r_total;
  if (T_app > n) return;
  const uint32_t *ptr = a.b.s_ptr + W.tb, *elem = a.b.s_elem + W.tb;
  uint32_t *dest = a.b.s_dest + W.tb;
  const uint32_t lo = r.b > W.r_total ? r.b : W.r_total;
  for (uint32_t t = lo + threadIdx.x; t < r.e && t < T_app; t += kBlock) {
    const uint32_t p = ptr[t];
    const uint32_t e = p < T_app ? elem[p] : n;
    if (e < n) dest[e] = W.n_left + (t - W.r_total);
  }
}

__global__ __launch_bounds__(kBlock) void k_tbw_scatter(TriWideArgs a) {
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  const uint32_t n = W.n, tb = W.tb;
  const uint32_t *dest = a.b.s_dest + tb;
  for (uint32_t k = r.b + threadIdx.x; k < r.e; k += kBlock) {
    const uint32_t d = dest[k];
    if (d >= n) continue;
    for (int c = 0; c < 6; ++c) a.b.box2[6 * (size_t)(tb + d) + c] = a.b.box[6 * (size_t)(tb + k) + c];
    a.b.ref2[tb + d] = a.b.ref[tb + k];
  }
}

__global__ __launch_bounds__(kBlock) void k_tbw_copy(TriWideArgs a) {
  const ChunkRange r = chunk_range(a, blockIdx.x);
  const TriWide &W = a.w[r.w];
  if (W.leaf) return;
  for (uint32_t k = r.b + threadIdx.x; k < r.e; k += kBlock) {
    const size_t p = W.tb + k;
    for (int c = 0; c < 6; ++c) a.b.box[6 * p + c] = a.b.box2[6 * p + c];
    a.b.ref[p] = a.b.ref2[p];
  }
}

__global__ void k_tbw_emit(TriWideArgs a) {
  if (threadIdx.x != 0) return;
  const TriWide &W = a.w[blockIdx.x];
  if (W.leaf) return;
  emit_children(a.b, W.id, W.axis, W.tb, W.tb + W.n_left, W.tb + W.n - W.n_right, W.tb + W.n, info_smask(W.info),
                info_depth(W.info));
}

}  // namespace

size_t tri_build_bins_bytes() { return sizeof(Bins); }

hipError_t launch_tri_wide_setup(const TriWideArgs &a, hipStream_t stream) {
  if (a.b.list_n == 0) return hipSuccess;
  if (a.chunk == 0 || a.chunk % kBlock != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tbw_setup, dim3(1), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_wide_level(const TriWideArgs &a, uint32_t jump_rounds, hipStream_t stream) {
  const uint32_t nw = a.b.list_n, nc = a.n_chunks;
  if (nw == 0 || nc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tbw_init, dim3(nw, 3), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_box, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_bins, dim3(nc, 3), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_split, dim3(nw), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_class, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_scan, dim3(nw), dim3(kBlock), 0, stream, a, 0);
  hipLaunchKernelGGL(k_tbw_prefix, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_K, dim3(nw), dim3(64), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_order, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_xcount, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_scan, dim3(nw), dim3(kBlock), 0, stream, a, 1);
  hipLaunchKernelGGL(k_tbw_exam, dim3(nc), dim3(kBlock), 0, stream, a);
  for (uint32_t i = 0; i < jump_rounds; ++i) hipLaunchKernelGGL(k_tbw_jump, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_mdest, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_scatter, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_copy, dim3(nc), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(k_tbw_emit, dim3(nw), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace skw
