// Device build of the ternary triangle tree (TriangleAccelerator::TriangleAccelerator, Node::Node
// and _DetermineBestSplit; Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:22-147, 505-729),
// node for node the tree of the host builder (host/builders.cpp), level by level:
//
//  * k_tb_level: one workgroup per node between the small-node size (at most 64) and the wide-node
//    size (larger nodes take the multi-workgroup kernels of tri_build_wide.hip, the same steps).
//    The node box is the fold of its triangles' boxes in array order (min / max keep the first of equal values, which
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
//  * k_tb_subtree: one wave per node of at most 64 triangles builds that node's whole subtree
//    depth first with the triangles' boxes in LDS: the same box fold, split search (a lane per
//    candidate bin position, every candidate evaluated directly, the first least cost over the
//    wave) and the sequential partition loop on a permutation of the node's positions.
// Children are allocated as contiguous id blocks; the host lays the tree out depth first as the
// host builder does.
// Every translation unit is compiled with -ffp-contract=off; f32 divisions are correctly rounded.
#include <hip/hip_runtime.h>

#include "tri_build_device.hpp"

namespace skw {
namespace {

using namespace tb;

// ------------------------------------------------------------------ one wave per small subtree

constexpr uint32_t kWaveMax = 64;    // triangles a subtree wave holds (one per lane)
constexpr uint32_t kStackMax = 416;  // DFS entries: at most 2 per level below the subtree's root, depth <= 200

struct WaveLds {
  float bx[kWaveMax][6];
  uint32_t ref[kWaveMax];
  uint32_t lbin[kWaveMax], rbin[kWaveMax];
  uint32_t cls[kWaveMax], perm[kWaveMax];
  uint32_t st_id[kStackMax], st_rng[kStackMax];
  uint32_t misc[4];  // stack size, stop flag, partition mid, partition right
};

// stack entry: local range [lo, hi), split-axis mask, depth
__device__ __forceinline__ uint32_t pack_rng(uint32_t lo, uint32_t hi, uint32_t smask, uint32_t depth) {
  return lo | (hi << 7) | (smask << 14) | (depth << 17);
}

__global__ __launch_bounds__(64) void k_tb_subtree(TriBuildArgs a) {
  __shared__ WaveLds W;
  const uint32_t lane = threadIdx.x;
  const uint32_t root = a.list[blockIdx.x];
  const uint32_t base = a.nodes[root].tb, n0 = a.nodes[root].te - base, rinfo = a.nodes[root].info;
  if (n0 > kWaveMax) {  // listed by mistake: reported, nothing written
    if (lane == 0) atomicOr(a.overflow, 2u);
    return;
  }
  if (lane < n0) {
    for (int c = 0; c < 6; ++c) W.bx[lane][c] = a.box[6 * (size_t)(base + lane) + c];
    W.ref[lane] = a.ref[base + lane];
  }
  if (lane == 0) {
    W.st_id[0] = root;
    W.st_rng[0] = pack_rng(0, n0, info_smask(rinfo), info_depth(rinfo));
    W.misc[0] = 1;
    W.misc[1] = 0;
  }
  __syncthreads();
  while (true) {
    const uint32_t sp = W.misc[0];
    if (sp == 0 || W.misc[1] != 0) break;
    const uint32_t id = W.st_id[sp - 1], rng = W.st_rng[sp - 1];
    __syncthreads();
    if (lane == 0) W.misc[0] = sp - 1;
    const uint32_t lo = rng & 127u, hi = (rng >> 7) & 127u, smask = (rng >> 14) & 7u, depth = rng >> 17;
    const uint32_t nn = hi - lo;
    // node box: the fold in array order (every lane)
    float mn[3] = {__int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000)};
    float mx[3] = {-mn[0], -mn[1], -mn[2]};
    for (uint32_t p = lo; p < hi; ++p)
      for (int c = 0; c < 3; ++c) {
        mn[c] = fmin_std(mn[c], W.bx[p][c]);
        mx[c] = fmax_std(mx[c], W.bx[p][3 + c]);
      }
    TriBuildNode &nd = a.nodes[id];
    if (lane == 0) {
      nd.tb = base + lo;
      nd.te = base + hi;
      for (int c = 0; c < 3; ++c) {
        nd.bmin[c] = mn[c];
        nd.bmax[c] = mx[c];
      }
    }
    if (leaf_by_rule(nn, smask, depth)) {
      if (lane == 0) make_leaf(a, id, smask, depth);
      __syncthreads();
      continue;
    }
    const double inv_area = 1.0 / (double)area_f(mn, mx);
    const uint32_t T = num_tries(nn);
    double axis_cost[3], axis_pos[3];
    for (int axis = 0; axis < 3; ++axis) {
      axis_cost[axis] = __builtin_huge_val();
      axis_pos[axis] = 0.0;
      if (smask & (1u << axis)) continue;
      const double coef = bin_coef(T, mn[axis], mx[axis]);
      const float flo = mn[axis];
      if (lane < nn) {
        uint32_t lbg, re;
        bins_of(W.bx[lo + lane], axis, flo, coef, T, lbg, re);
        W.lbin[lane] = lbg;
        W.rbin[lane] = re;
      }
      __syncthreads();
      // candidates lane and lane + 64: left = bins 0..i, right = bins i..T-1, middle = straddling i
      double my_cost = __builtin_huge_val();
      uint32_t my_i = 0xFFFFFFFFu;
      for (uint32_t i = lane; i < T; i += kWaveMax) {
        float lmn[3], lmx[3], mmn[3], mmx[3], rmn[3], rmx[3];
        for (int c = 0; c < 3; ++c) {
          lmn[c] = mmn[c] = rmn[c] = __int_as_float(0x7F800000);
          lmx[c] = mmx[c] = rmx[c] = __int_as_float(0xFF800000);
        }
        uint32_t ln = 0, mc = 0, rn = 0;
        for (uint32_t k = 0; k < nn; ++k) {
          const uint32_t lbg = W.lbin[k], re = W.rbin[k];
          const float *bx = W.bx[lo + k];
          const bool isl = lbg <= i && lbg < T, isr = re > i, ism = re <= i && i < lbg;
          ln += isl ? 1u : 0u;
          rn += isr ? 1u : 0u;
          mc += ism ? 1u : 0u;
          for (int c = 0; c < 3; ++c) {
            if (isl) {
              lmn[c] = fmin_std(lmn[c], bx[c]);
              lmx[c] = fmax_std(lmx[c], bx[3 + c]);
            }
            if (isr) {
              rmn[c] = fmin_std(rmn[c], bx[c]);
              rmx[c] = fmax_std(rmx[c], bx[3 + c]);
            }
            if (ism) {
              mmn[c] = fmin_std(mmn[c], bx[c]);
              mmx[c] = fmax_std(mmx[c], bx[3 + c]);
            }
          }
        }
        double dl[3], dL[3], dm[3], dM[3], dr[3], dR[3];
        for (int c = 0; c < 3; ++c) {
          dl[c] = lmn[c];
          dL[c] = lmx[c];
          dm[c] = mmn[c];
          dM[c] = mmx[c];
          dr[c] = rmn[c];
          dR[c] = rmx[c];
        }
        const double p1 = dmin_std(1.0, area_d(dl, dL) * inv_area);
        const double p2 = dmin_std(1.0, area_d(dm, dM) * inv_area);
        const double p3 = dmin_std(1.0, area_d(dr, dR) * inv_area);
        double cost = (double)ln * p1 + (double)mc * p2 + (double)rn * p3;
        if (ln == 0 || rn == 0) cost *= 0.8;
        if (cost < my_cost) {
          my_cost = cost;
          my_i = i;
        }
      }
      // the first candidate of least cost
      for (int off = 32; off >= 1; off >>= 1) {
        const double oc = __shfl_xor(my_cost, off);
        const uint32_t oi = __shfl_xor(my_i, off);
        if (oc < my_cost || (oc == my_cost && oi < my_i)) {
          my_cost = oc;
          my_i = oi;
        }
      }
      axis_cost[axis] = my_cost;
      if (my_i != kNone) axis_pos[axis] = split_pos(my_i, T, mn[axis], mx[axis]);
      __syncthreads();
    }
    const uint32_t axis = pick_axis(axis_cost[0], axis_cost[1], axis_cost[2]);
    const double split = axis_pos[axis];
    // the sequential three-way partition (TriangleAccelerator.cpp:650-701) on a permutation of the
    // node's positions, then applied to the boxes
    if (lane < nn) {
      const float *bx = W.bx[lo + lane];
      W.cls[lane] = class_of(bx, axis, split);
      W.perm[lane] = lane;
    }
    __syncthreads();
    if (lane == 0) {
      uint32_t mid = 0, right = nn, i = 0;
      while (i < right) {
        const uint32_t c = W.cls[W.perm[i]];
        if (c == 0u) {
          const uint32_t t = W.perm[mid];
          W.perm[mid] = W.perm[i];
          W.perm[i] = t;
          ++mid;
          ++i;
        } else if (c == 2u) {
          --right;
          const uint32_t t = W.perm[i];
          W.perm[i] = W.perm[right];
          W.perm[right] = t;
        } else {
          ++i;
        }
      }
      W.misc[2] = mid;
      W.misc[3] = right;
    }
    __syncthreads();
    float v[6];
    uint32_t r = 0;
    if (lane < nn) {
      const uint32_t src = lo + W.perm[lane];
      for (int c = 0; c < 6; ++c) v[c] = W.bx[src][c];
      r = W.ref[src];
    }
    __syncthreads();
    if (lane < nn) {
      for (int c = 0; c < 6; ++c) W.bx[lo + lane][c] = v[c];
      W.ref[lo + lane] = r;
    }
    if (lane == 0) {
      const uint32_t mid = lo + W.misc[2], right = lo + W.misc[3];
      const uint32_t rb[3] = {lo, mid, right}, re[3] = {mid, right, hi};
      uint32_t cmask = 0, nch = 0;
      for (int c = 0; c < 3; ++c)
        if (re[c] > rb[c]) {
          cmask |= 1u << c;
          ++nch;
        }
      const uint32_t first = atomicAdd(a.node_count, nch);
      uint32_t top = W.misc[0];
      if (first + nch > a.node_cap || top + nch > kStackMax) {
        atomicOr(a.overflow, 1u);
        W.misc[1] = 1;
      } else {
        nd.first_child = first;
        nd.info = (axis + 1u) | (cmask << 2) | (smask << 5) | (depth << 16);
        uint32_t k = 0;
        for (int c = 0; c < 3; ++c) {
          if (re[c] <= rb[c]) continue;
          const uint32_t cm = c == 1 ? (smask | (1u << axis)) : smask;
          W.st_id[top] = first + k++;
          W.st_rng[top] = pack_rng(rb[c], re[c], cm, depth + 1u);
          ++top;
        }
        W.misc[0] = top;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (lane < n0) {
    for (int c = 0; c < 6; ++c) a.box[6 * (size_t)(base + lane) + c] = W.bx[lane][c];
    a.ref[base + lane] = W.ref[lane];
  }
}

// ------------------------------------------------------------------ one workgroup per larger node

struct Lds {
  Bins B;
  double cost[kMaxTries];
  unsigned long long key[6];
  unsigned long long wsum[kBlock / 64];
  double axis_cost[3], axis_pos[3];
  uint32_t misc[8];
};

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
    BoxKeys k;
    k.clear();
    for (uint32_t p = tb + tid; p < te; p += kBlock) k.add(a.box + 6 * (size_t)p, p);
    keys_merge(L.key, k);
  }
  __syncthreads();
  float mn[3], mx[3];
  keys_decode(L.key, a.box, mn, mx);
  if (tid == 0)
    for (int c = 0; c < 3; ++c) {
      nd.bmin[c] = mn[c];
      nd.bmax[c] = mx[c];
    }
  if (leaf_by_rule(n, smask, depth)) {
    if (tid == 0) make_leaf(a, id, smask, depth);
    return;
  }
  const double inv_area = 1.0 / (double)area_f(mn, mx);
  const uint32_t T = num_tries(n);

  // ---- split search per axis
  for (int axis = 0; axis < 3; ++axis) {
    if (smask & (1u << axis)) {
      if (tid == 0) {
        L.axis_cost[axis] = __builtin_huge_val();
        L.axis_pos[axis] = 0.0;
      }
      continue;
    }
    bins_clear(L.B, tid, kBlock);
    __syncthreads();
    const double coef = bin_coef(T, mn[axis], mx[axis]);
    const float lo = mn[axis];
    for (uint32_t p = tb + tid; p < te; p += kBlock) {
      const float *bx = a.box + 6 * (size_t)p;
      uint32_t lbg, re;
      bins_of(bx, axis, lo, coef, T, lbg, re);
      bins_add(L.B, bx, lbg, re, T);
    }
    __syncthreads();
    for (uint32_t i = tid; i < T; i += kBlock) L.cost[i] = bins_cost(L.B, i, T, inv_area);
    __syncthreads();
    if (tid == 0) first_min(L.cost, T, mn[axis], mx[axis], L.axis_cost[axis], L.axis_pos[axis]);
    __syncthreads();
  }
  __syncthreads();  // a masked axis's entries are written without a barrier of their own
  const uint32_t axis = pick_axis(L.axis_cost[0], L.axis_cost[1], L.axis_cost[2]);
  const double split = L.axis_pos[axis];

  // ---- exact parallel form of the three-way partition (positions relative to tb)
  uint32_t *cls = a.s_cls + tb, *rf = a.s_rf + tb, *nb = a.s_nb + tb, *G = a.s_G + tb, *Rpos = a.s_Rpos + tb;
  uint32_t *order = a.s_order + tb, *elem = a.s_elem + tb, *ptr = a.s_ptr + tb, *dest = a.s_dest + tb;
  // classes; right elements before each front position; non-right ones before each back position
  unsigned long long carry_r = 0, carry_b = 0, tot = 0;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t k = base + tid;
    uint32_t ck = 3u;
    if (k < n) {
      ck = class_of(a.box + 6 * (size_t)(tb + k), axis, split);
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

// ------------------------------------------------------------------ world-space triangles

// Transform::operator()(Point3D<float>) (Source/RayLibs/Math/Transform.h:223-238) as the host
// builder computes it: double products of the float point, float results, the w division in
// double only when w != 1.
__device__ __forceinline__ void xform_point_f(const double m[16], const float *p, float *out) {
  const double x = p[0], y = p[1], z = p[2];
  const float xp = (float)(m[0] * x + m[1] * y + m[2] * z + m[3]);
  const float yp = (float)(m[4] * x + m[5] * y + m[6] * z + m[7]);
  const float zp = (float)(m[8] * x + m[9] * y + m[10] * z + m[11]);
  const float wp = (float)(m[12] * x + m[13] * y + m[14] * z + m[15]);
  if (wp == 1.0f) {
    out[0] = xp;
    out[1] = yp;
    out[2] = zp;
    return;
  }
  const double inv = 1.0 / (double)wp;
  out[0] = (float)(xp * inv);
  out[1] = (float)(yp * inv);
  out[2] = (float)(zp * inv);
}

__global__ __launch_bounds__(256) void k_tb_prep(TriPrepArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  // the mesh holding triangle i: the last mesh whose first triangle is <= i
  uint32_t lo = 0, hi = a.n_meshes;  // goff[lo] <= i < goff[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (a.goff[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t m = lo;
  const uint64_t j = i - a.goff[m];
  const DevMesh mesh = a.meshes[m];
  const double *m2w = a.prims[a.user[m]].m2w;
  float mn[3] = {__int_as_float(0x7F800000), __int_as_float(0x7F800000), __int_as_float(0x7F800000)};
  float mx[3] = {-mn[0], -mn[1], -mn[2]};
  for (int k = 0; k < 3; ++k) {
    const uint32_t vi = a.tris[3 * (mesh.tri_off + j) + k];
    float w[3];
    xform_point_f(m2w, a.verts + 3 * (mesh.vert_off + vi), w);
    for (int c = 0; c < 3; ++c) {
      a.tv[9 * i + 3 * k + c] = w[c];
      mn[c] = fmin_std(mn[c], w[c]);
      mx[c] = fmax_std(mx[c], w[c]);
    }
  }
  bool fin = true;
  for (int c = 0; c < 3; ++c) {
    a.box[6 * i + c] = mn[c];
    a.box[6 * i + 3 + c] = mx[c];
    fin = fin && isfinite(mn[c]) && isfinite(mx[c]);
  }
  a.tprim[i] = a.user[m];
  a.ttri[i] = (uint32_t)j;
  a.ref[i] = (uint32_t)i;
  if (!fin) atomicOr(a.flags, 1u);
}

// ------------------------------------------------------------------ depth-first layout

__global__ __launch_bounds__(256) void k_tl_hist(TriLayoutArgs a) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.count) return;
  const TriBuildNode nd = a.hn[v];
  const uint32_t depth = info_depth(nd.info);
  bool bad = depth >= kTriLayoutDepths;
  if (nd.first_child == kNone) {
    atomicAdd(a.leaves, 1u);
    bad = bad || nd.te < nd.tb || nd.te > a.n;
  } else {
    // children are allocated after their parent, as one block inside the records
    const uint32_t nch = __builtin_popcount((nd.info >> 2) & 7u);
    bad = bad || nch == 0 || nd.first_child <= v || (uint64_t)nd.first_child + nch > a.count;
  }
  if (bad) {
    atomicOr(a.flags, 2u);
    return;
  }
  atomicAdd(&a.hist[depth], 1u);
}

__global__ __launch_bounds__(256) void k_tl_scatter(TriLayoutArgs a) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.count) return;
  const uint32_t depth = info_depth(a.hn[v].info);
  a.by_depth[a.level_off[depth] + atomicAdd(&a.fill[depth], 1u)] = v;
  if (v == 0) {
    a.idx[0] = 0;
    a.cval[0] = 0;
  }
}

__global__ __launch_bounds__(256) void k_tl_size(TriLayoutArgs a) {
  const uint32_t t = a.level_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.level_end) return;
  const uint32_t v = a.by_depth[t];
  const TriBuildNode nd = a.hn[v];
  uint32_t sz = 1;
  if (nd.first_child != kNone) {
    const uint32_t nch = __builtin_popcount((nd.info >> 2) & 7u);
    for (uint32_t k = 0; k < nch; ++k) sz += a.size[nd.first_child + k];
  }
  a.size[v] = sz;
}

// Node v's children block starts at 1 + cval(v), cval(v) = children of the nodes placed before v;
// the first child follows v (cval + v's children), each next child follows the previous child's
// subtree (cval + its size - 1: every node of a subtree but its top is a child inside it).
__global__ __launch_bounds__(256) void k_tl_place(TriLayoutArgs a) {
  const uint32_t t = a.level_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.level_end) return;
  const uint32_t v = a.by_depth[t];
  const TriBuildNode nd = a.hn[v];
  DevNode out;
  for (int c = 0; c < 3; ++c) {
    out.bmin[c] = nd.bmin[c];
    out.bmax[c] = nd.bmax[c];
  }
  if (nd.first_child == kNone) {
    out.a = nd.tb;
    out.b = (nd.te - nd.tb) << 3;
  } else {
    const uint32_t mask = (nd.info >> 2) & 7u, nch = __builtin_popcount(mask);
    const uint32_t bs = 1u + a.cval[v];
    out.a = bs;
    out.b = (nd.info & 3u) | (mask << 2);
    uint32_t cv = a.cval[v] + nch;
    for (uint32_t k = 0; k < nch; ++k) {
      const uint32_t c = nd.first_child + k;
      a.idx[c] = bs + k;
      a.cval[c] = cv;
      cv += a.size[c] - 1u;
    }
  }
  a.nodes[a.idx[v]] = out;
}

__global__ __launch_bounds__(256) void k_tl_tris(TriLayoutArgs a) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= a.n) return;
  const uint32_t id = a.ref[pos];
  DevTri t;
  const float *v = a.tv + 9 * (size_t)id;
  for (int c = 0; c < 3; ++c) {
    t.v0[c] = v[c];
    t.v1[c] = v[3 + c];
    t.v2[c] = v[6 + c];
  }
  t.prim = a.tprim[id];
  t.tri = a.ttri[id];
  t.pad = 0;
  a.tris[pos] = t;
}

}  // namespace

hipError_t launch_tri_prep(const TriPrepArgs &a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tb_prep, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_layout_hist(const TriLayoutArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(k_tl_hist, dim3((a.count + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_layout_scatter(const TriLayoutArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(k_tl_scatter, dim3((a.count + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_layout_size(const TriLayoutArgs &a, hipStream_t stream) {
  if (a.level_end <= a.level_begin) return hipSuccess;
  hipLaunchKernelGGL(k_tl_size, dim3((a.level_end - a.level_begin + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_layout_place(const TriLayoutArgs &a, hipStream_t stream) {
  if (a.level_end <= a.level_begin) return hipSuccess;
  hipLaunchKernelGGL(k_tl_place, dim3((a.level_end - a.level_begin + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_layout_tris(const TriLayoutArgs &a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tl_tris, dim3((a.n + 255) / 256), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_build_level(const TriBuildArgs &a, hipStream_t stream) {
  if (a.list_n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tb_level, dim3(a.list_n), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_tri_build_small(const TriBuildArgs &a, hipStream_t stream) {
  if (a.list_n == 0) return hipSuccess;
  if (a.small_max > kWaveMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tb_subtree, dim3(a.list_n), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace skw
