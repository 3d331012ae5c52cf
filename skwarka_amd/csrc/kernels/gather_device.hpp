// Device-side photon-map gather, shared by the gather kernels (gather.hip) and the multi-scene
// batch kernel (batch.hip): k-nearest / nearest lookups in the photon kd-tree with the
// PhotonFilter, and the density estimates of the photon integrator.
//
// Lookup order follows KDTree::_Lookup (Source/RayLibs/Raytracer/Core/KDTree.h:535-576): the node's
// own photon first (strict d^2 < max), then the near child (q[axis] <= split: left), then the far
// child if the plane distance^2 < the radius^2 current at that moment. The recursion becomes a
// per-lane stack of parent indices in LDS ([depth][lane]); the far-side test is made when the
// entry is popped, i.e. after the near subtree, as in the recursive form.
// The bounded heap is NearestPointsProc (KDTree.h:331-387): fill unordered up to k, then
// make_heap; afterwards pop_heap / replace last / push_heap, ordered by (d^2, tree position)
// (KDTree.h:174-177). The heap routines are written out step for step after libstdc++'s
// std::make_heap / __adjust_heap / __push_heap, so the final array order, and the order in which
// the estimate sums the photons, is the same as a libstdc++ build of the reference.
// Every translation unit including this is compiled with -ffp-contract=off; each gets its own
// copy of these functions (internal linkage).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace skw {
namespace gath {
namespace {

constexpr int kWave = 64;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kPi = 3.14159265358979323846;
constexpr double kCosThreshold = 0.87;  // PhotonLTEIntegrator.cpp:27

// Lookup stacks: LDS [depth][lane] of one wave, or global memory [depth][query].
template <int S>
struct LdsStack {
  uint32_t *base;  // &lds[lane]
  __device__ __forceinline__ uint32_t get(uint32_t i) const { return base[i * kWave]; }
  __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { base[i * kWave] = v; }
};
struct GlobalStack {
  uint32_t *base;
  uint64_t stride;
  __device__ __forceinline__ uint32_t get(uint32_t i) const { return base[i * stride]; }
  __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { base[i * stride] = v; }
};

struct Heap {
  uint32_t *pos;
  double *d2;
  uint64_t js;  // element stride
  __device__ __forceinline__ uint32_t p(int64_t i) const { return pos[i * js]; }
  __device__ __forceinline__ double d(int64_t i) const { return d2[i * js]; }
  __device__ __forceinline__ void set(int64_t i, uint32_t pp, double dd) const {
    pos[i * js] = pp;
    d2[i * js] = dd;
  }
};

__device__ __forceinline__ bool np_less(double da, uint32_t pa, double db, uint32_t pb) {
  return da == db ? (pa < pb) : (da < db);
}

// libstdc++ __push_heap
__device__ void push_heap_impl(const Heap &h, int64_t hole, int64_t top, uint32_t vp, double vd) {
  int64_t parent = (hole - 1) / 2;
  while (hole > top && np_less(h.d(parent), h.p(parent), vd, vp)) {
    h.set(hole, h.p(parent), h.d(parent));
    hole = parent;
    parent = (hole - 1) / 2;
  }
  h.set(hole, vp, vd);
}

// libstdc++ __adjust_heap
__device__ void adjust_heap(const Heap &h, int64_t hole, int64_t len, uint32_t vp, double vd) {
  const int64_t top = hole;
  int64_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (np_less(h.d(child), h.p(child), h.d(child - 1), h.p(child - 1))) child--;
    h.set(hole, h.p(child), h.d(child));
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    h.set(hole, h.p(child - 1), h.d(child - 1));
    hole = child - 1;
  }
  push_heap_impl(h, hole, top, vp, vd);
}

__device__ void make_heap(const Heap &h, int64_t len) {
  if (len < 2) return;
  int64_t parent = (len - 2) / 2;
  while (true) {
    uint32_t vp = h.p(parent);
    double vd = h.d(parent);
    adjust_heap(h, parent, len, vp, vd);
    if (parent == 0) return;
    parent--;
  }
}

// pop_heap(first, first + len): moves the top to len - 1 and re-heaps [0, len - 1).
__device__ void pop_heap(const Heap &h, int64_t len) {
  if (len > 1) {
    int64_t last = len - 1;
    uint32_t vp = h.p(last);
    double vd = h.d(last);
    h.set(last, h.p(0), h.d(0));
    adjust_heap(h, 0, last, vp, vd);
  }
}

__device__ void push_heap(const Heap &h, int64_t len) {
  uint32_t vp = h.p(len - 1);
  double vd = h.d(len - 1);
  push_heap_impl(h, len - 1, 0, vp, vd);
}

struct Query {
  double q0, q1, q2;
  double n0, n1, n2;
  double sin_t;
};

// Query coordinate on a split axis, selected as a value (a conditional over the struct's members
// would select their addresses and move the query into scratch memory).
__device__ __forceinline__ double axis_coord(const Query &q, uint32_t axis) {
  const double x = q.q0, y = q.q1, z = q.q2;
  return axis == 0 ? x : (axis == 1 ? y : z);
}

// PhotonFilter::operator() (PhotonInternalTypes.h:128-134); to = photon - point, already formed
// by the distance computation with the same subtraction.
__device__ __forceinline__ bool photon_filter(const GatherArgs &a, const Query &q, uint32_t pos, double dx, double dy,
                                              double dz) {
  uint32_t code = a.photons[pos].codes >> 16;
  const double *nv = a.cdir + 3 * (size_t)code;
  double c = nv[0] * q.n0 + nv[1] * q.n1 + nv[2] * q.n2;
  if (!(c > kCosThreshold)) return false;
  double along = dx * q.n0 + dy * q.n1 + dz * q.n2;
  return fabs(along) < q.sin_t * sqrt(dx * dx + dy * dy + dz * dz);
}

template <int MODE, class Stack>
__device__ void lookup(const GatherArgs &a, const Query &q, double &max_d2, const Stack &stk, uint32_t &size,
                       int64_t &nearest, const Heap &heap) {
  if (a.n_nodes == 0) return;
  const uint32_t k = a.k;
  uint32_t sp = 0;
  uint32_t cur = 0;
  bool have = true;
  while (true) {
    if (!have) {
      // Pop the deepest pending far side; visit it if the plane is still inside the radius.
      bool next = false;
      while (sp > 0) {
        uint32_t par = stk.get(--sp);
        float4 nd = reinterpret_cast<const float4 *>(a.nodes)[par];
        uint32_t w = __float_as_uint(nd.w);
        uint32_t axis = w & 3u;
        double qa = axis_coord(q, axis);
        double split = axis == 0 ? (double)nd.x : (axis == 1 ? (double)nd.y : (double)nd.z);
        double pd2 = (qa - split) * (qa - split);
        if (pd2 < max_d2) {
          cur = (qa <= split) ? (w >> 3) : par + 1;
          next = true;
          break;
        }
      }
      if (!next) break;
    }
    have = false;
    float4 nd = reinterpret_cast<const float4 *>(a.nodes)[cur];
    double dx = (double)nd.x - q.q0, dy = (double)nd.y - q.q1, dz = (double)nd.z - q.q2;
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < max_d2) {
      bool pass = a.filter == 0 || photon_filter(a, q, cur, dx, dy, dz);
      if (pass) {
        if (MODE == 1) {
          nearest = cur;
          max_d2 = d2;
        } else if (size < k) {
          heap.set(size, cur, d2);
          ++size;
          if (size == k) {
            make_heap(heap, size);
            max_d2 = heap.d(0);
          }
        } else {
          pop_heap(heap, size);
          heap.set(k - 1, cur, d2);
          push_heap(heap, size);
          max_d2 = heap.d(0);
        }
      }
    }
    uint32_t w = __float_as_uint(nd.w);
    uint32_t axis = w & 3u;
    if (axis == 3u) continue;  // leaf
    double qa = axis_coord(q, axis);
    double split = axis == 0 ? (double)nd.x : (axis == 1 ? (double)nd.y : (double)nd.z);
    bool has_left = (w & 4u) != 0;
    uint32_t right = w >> 3;
    bool has_right = right < a.n_nodes;
    bool left_near = qa <= split;
    bool near_exists = left_near ? has_left : has_right;
    bool far_exists = left_near ? has_right : has_left;
    if (far_exists) stk.set(sp++, cur);
    if (near_exists) {
      cur = left_near ? cur + 1 : right;
      have = true;
    }
  }
}

// Cone-filtered irradiance over the set (Jensen's cone filter; not an estimator of the
// reference): photon weight 1 - d / (kc * r_max), only photons arriving from the normal's side,
// normalised by (1 - 2 / (3 kc)) * pi * paths * r_max^2, with r_max^2 = the set's largest d^2, or
// r^2 when the set is not full.
__device__ void cone_estimate(const GatherArgs &a, const Query &q, uint64_t i, const Heap &heap, uint32_t size) {
  float *out = a.e_ext + 3 * i;
  if (size == 0) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  double mx = 0.0;
  for (uint32_t j = 0; j < size; ++j) mx = (mx < heap.d(j)) ? heap.d(j) : mx;
  const double rmax2 = (size < a.k || mx == 0.0) ? a.r * a.r : mx;
  const double rmax = sqrt(rmax2);
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (uint32_t j = 0; j < size; ++j) {
    float4 ph = reinterpret_cast<const float4 *>(a.photons)[heap.p(j)];
    const double *dv = a.cdir + 3 * (size_t)(__float_as_uint(ph.w) & 0xFFFFu);
    if (!(dv[0] * q.n0 + dv[1] * q.n1 + dv[2] * q.n2 > 0.0)) continue;
    const double wgt = 1.0 - sqrt(heap.d(j)) / (a.kc * rmax);
    e0 += wgt * (double)ph.x;
    e1 += wgt * (double)ph.y;
    e2 += wgt * (double)ph.z;
  }
  const double norm = (1.0 - 2.0 / (3.0 * a.kc)) * kPi * a.paths * rmax2;
  out[0] = (float)(e0 / norm);
  out[1] = (float)(e1 / norm);
  out[2] = (float)(e2 / norm);
}

// PhotonLTEIntegrator::_LookupCausticRadiance / _PhotonKernel (PhotonLTEIntegrator.cpp:579-636) for
// a Lambertian surface of albedo rho: Simpson kernel 3/pi (1 - d^2 / m)^2 with m the running
// maximum of d^2 in the set's array order, BSDF value rho/pi when photon direction and wo lie on
// the same side of the normal; then / (paths * max_d2) with the disc estimate's n / (n - 0.5) rule.
__device__ void caustic_estimate(const GatherArgs &a, const Query &q, uint64_t i, const Heap &heap, uint32_t size) {
  double *out = a.caustic + 3 * i;
  if (size == 0) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  const double f = a.rho * kInvPi;
  const double *wv = a.wo + 3 * i;
  const double wn = wv[0] * q.n0 + wv[1] * q.n1 + wv[2] * q.n2;
  double mx = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
  for (uint32_t j = 0; j < size; ++j) {
    const uint32_t pos = heap.p(j);
    float4 nd = reinterpret_cast<const float4 *>(a.nodes)[pos];
    double dx = (double)nd.x - q.q0, dy = (double)nd.y - q.q1, dz = (double)nd.z - q.q2;
    double t = dx * dx + dy * dy + dz * dz;
    if (t > mx) mx = t;
    double tmp = 1.0 - t / mx;
    double kernel = 3.0 * kInvPi * tmp * tmp;
    float4 ph = reinterpret_cast<const float4 *>(a.photons)[pos];
    const double *dv = a.cdir + 3 * (size_t)(__float_as_uint(ph.w) & 0xFFFFu);
    double pn = dv[0] * q.n0 + dv[1] * q.n1 + dv[2] * q.n2;
    double fv = (pn * wn > 0) ? f : 0.0;
    r0 += fv * (double)ph.x * kernel;
    r1 += fv * (double)ph.y * kernel;
    r2 += fv * (double)ph.z * kernel;
  }
  if (size < a.k || mx == 0.0)
    mx = a.r * a.r;
  else
    mx *= (double)size / ((double)size - 0.5);
  const double inv = 1.0 / (a.paths * mx);
  out[0] = r0 * inv;
  out[1] = r1 * inv;
  out[2] = r2 * inv;
}

template <int MODE, class Stack>
__device__ void gather_one(const GatherArgs &a, uint64_t i, const Stack &stk) {
  Query q;
  const double *pp = a.points + 3 * i;
  q.q0 = pp[0];
  q.q1 = pp[1];
  q.q2 = pp[2];
  if (a.normals) {
    const double *nn = a.normals + 3 * i;
    q.n0 = nn[0];
    q.n1 = nn[1];
    q.n2 = nn[2];
  } else {
    q.n0 = q.n1 = q.n2 = 0.0;
  }
  q.sin_t = a.sin_t;
  double max_d2 = a.r * a.r;
  uint32_t size = 0;
  int64_t nearest = -1;
  Heap heap{a.heap_pos + i * a.q_stride, a.heap_d2 + i * a.q_stride, a.j_stride};
  if (MODE == 1) {
    lookup<1>(a, q, max_d2, stk, size, nearest, heap);
    a.nearest[i] = nearest < 0 ? kMiss : a.perm[nearest];
    return;
  }
  if (a.k > 0) lookup<0>(a, q, max_d2, stk, size, nearest, heap);
  if (a.count) a.count[i] = size;
  if (a.ids || a.d2) {
    for (uint32_t j = 0; j < size; ++j) {
      uint32_t pos = heap.p(j);
      if (a.ids) a.ids[i * a.k + j] = a.perm[pos];
      if (a.d2) a.d2[i * a.k + j] = heap.d(j);
    }
  }
  if (a.estimate == 1) {
    cone_estimate(a, q, i, heap, size);
    return;
  }
  if (a.estimate == 2) {
    caustic_estimate(a, q, i, heap, size);
    return;
  }
  if (!a.e_ext) return;
  // PhotonLTEIntegrator.cpp:139-168
  double ext0 = 0.0, ext1 = 0.0, ext2 = 0.0, in0 = 0.0, in1 = 0.0, in2 = 0.0;
  double mx = 0.0;
  for (uint32_t j = 0; j < size; ++j) {
    uint32_t pos = heap.p(j);
    double t = heap.d(j);
    if (t > mx) mx = t;
    float4 ph = reinterpret_cast<const float4 *>(a.photons)[pos];
    uint32_t dcode = __float_as_uint(ph.w) & 0xFFFFu;
    const double *dv = a.cdir + 3 * (size_t)dcode;
    double c = dv[0] * q.n0 + dv[1] * q.n1 + dv[2] * q.n2;
    if (c > 0.0) {
      ext0 += (double)ph.x;
      ext1 += (double)ph.y;
      ext2 += (double)ph.z;
    } else {
      in0 += (double)ph.x;
      in1 += (double)ph.y;
      in2 += (double)ph.z;
    }
  }
  float e[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (size > 0) {
    if (size < a.k || mx == 0.0)
      mx = a.r * a.r;
    else
      mx *= (double)size / ((double)size - 0.5);
    double inv = 1.0 / (kPi * a.paths * mx);
    e[0] = (float)(ext0 * inv);
    e[1] = (float)(ext1 * inv);
    e[2] = (float)(ext2 * inv);
    e[3] = (float)(in0 * inv);
    e[4] = (float)(in1 * inv);
    e[5] = (float)(in2 * inv);
  }
  for (int c = 0; c < 3; ++c) {
    a.e_ext[3 * i + c] = e[c];
    if (a.e_int) a.e_int[3 * i + c] = e[3 + c];
    if (a.radiance) a.radiance[3 * i + c] = (float)(a.rho * kInvPi * (double)e[c]);
  }
}


}  // namespace
}  // namespace gath
}  // namespace skw
