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
__device__ void push_heap_impl(Heap h, int64_t hole, int64_t top, uint32_t vp, double vd) {
  int64_t parent = (hole - 1) / 2;
  while (hole > top && np_less(h.d(parent), h.p(parent), vd, vp)) {
    h.set(hole, h.p(parent), h.d(parent));
    hole = parent;
    parent = (hole - 1) / 2;
  }
  h.set(hole, vp, vd);
}

// libstdc++ __adjust_heap
__device__ void adjust_heap(Heap h, int64_t hole, int64_t len, uint32_t vp, double vd) {
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

__device__ void make_heap(Heap h, int64_t len) {
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
__device__ void pop_heap(Heap h, int64_t len) {
  if (len > 1) {
    int64_t last = len - 1;
    uint32_t vp = h.p(last);
    double vd = h.d(last);
    h.set(last, h.p(0), h.d(0));
    adjust_heap(h, 0, last, vp, vd);
  }
}

__device__ void push_heap(Heap h, int64_t len) {
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
// by the distance computation with the same subtraction; codes = the photon's packed direction
// and normal ids.
__device__ __forceinline__ bool photon_filter(const GatherArgs &a, const Query &q, uint32_t codes, double dx,
                                              double dy, double dz) {
  const double *nv = a.cdir + 3 * (size_t)(codes >> 16);
  double c = nv[0] * q.n0 + nv[1] * q.n1 + nv[2] * q.n2;
  if (!(c > kCosThreshold)) return false;
  double along = dx * q.n0 + dy * q.n1 + dz * q.n2;
  return fabs(along) < q.sin_t * sqrt(dx * dx + dy * dy + dz * dz);
}

// photon_filter with the photon's decoded normal given (a nearest-lookup record's copy of the
// direction table's values): the same arithmetic.
__device__ __forceinline__ bool photon_filter_nv(const Query &q, double nv0, double nv1, double nv2, double dx,
                                                 double dy, double dz) {
  double c = nv0 * q.n0 + nv1 * q.n1 + nv2 * q.n2;
  if (!(c > kCosThreshold)) return false;
  double along = dx * q.n0 + dy * q.n1 + dz * q.n2;
  return fabs(along) < q.sin_t * sqrt(dx * dx + dy * dy + dz * dz);
}

// True when every normal in the box fails the PhotonFilter's test (decoded normal . query normal
// > 0.87): the dot product's upper bound over the box is below 0.87 by a margin far above the
// rounding of either computation.
__device__ __forceinline__ bool normals_fail(uint64_t box, const Query &q) {
  const double s = 1.0 / 127.0;
  const double lo0 = (double)(int8_t)(box & 0xFF) * s, lo1 = (double)(int8_t)((box >> 8) & 0xFF) * s,
               lo2 = (double)(int8_t)((box >> 16) & 0xFF) * s;
  const double hi0 = (double)(int8_t)((box >> 24) & 0xFF) * s, hi1 = (double)(int8_t)((box >> 32) & 0xFF) * s,
               hi2 = (double)(int8_t)((box >> 40) & 0xFF) * s;
  const double bound = fmax(q.n0 * lo0, q.n0 * hi0) + fmax(q.n1 * lo1, q.n1 * hi1) + fmax(q.n2 * lo2, q.n2 * hi2);
  return bound < kCosThreshold - 1e-9;
}

// Adds a photon's weight to the exitant (ext) or the incoming-side sums (PhotonLTEIntegrator.cpp:
// 150-153). Both sums are updated with value selects (the other one gains +0.0, which leaves a
// sum of non-negative weights unchanged bit for bit); an if/else over the sums would let the
// compiler select their addresses and move them to scratch memory.
__device__ __forceinline__ void add_split(bool ext, const float4 &ph, double &e0, double &e1, double &e2, double &i0,
                                          double &i1, double &i2) {
  const double x = ph.x, y = ph.y, z = ph.z;
  e0 += ext ? x : 0.0;
  e1 += ext ? y : 0.0;
  e2 += ext ? z : 0.0;
  i0 += ext ? 0.0 : x;
  i1 += ext ? 0.0 : y;
  i2 += ext ? 0.0 : z;
}

// The disc estimate's sums (PhotonLTEIntegrator.cpp:139-153) kept while the set is filled: until
// the set holds k photons it is an array in insertion order, which is the order the estimate
// sums it in, so adding each photon as it is inserted gives the same sums bit for bit. Once the
// set is full (make_heap reorders it), `full` is set and the sums are taken from the final array.
struct DiscSums {
  double ext0, ext1, ext2, in0, in1, in2, mx;
  bool full;
};

// Work counters of one lookup (the statistics instance only): nodes visited, photons inside the
// radius (those reach the PhotonFilter).
struct Counts {
  uint32_t visits, filtered;
};

// Visits node `cur` of a lookup (KDTree.h:545-575): tests its photon (strict d^2 < max_d2, then
// the PhotonFilter, then the bounded set), then picks the children. Returns true when the node
// has a near child (set in `cur`); `far` / `axis` / `split` describe a far child to visit later
// when `far_exists`.
// STORE = false: the set's slots are not written (the disc estimate's sums are kept instead); the
// visit that would fill the set sets `filled` and the caller restarts the lookup with STORE.
// STATS: counts visited nodes and photons that reach the PhotonFilter (inside the radius).
template <int MODE, bool SUMS, bool STORE, bool STATS>
__device__ __forceinline__ bool visit(const GatherArgs &a, const Query &q, uint32_t &cur, double &max_d2,
                                      uint32_t &size, int64_t &nearest, const Heap &heap, DiscSums &sums,
                                      bool &far_exists, uint32_t &far, uint32_t &axis, float &split, bool &filled,
                                      Counts &cnt) {
  const uint32_t k = a.k;
  // A nearest lookup over a map with records reads the node, the photon's normal and the subtree's
  // normal box from one 64-byte record; otherwise the node, its photon record and the normal box
  // are independent loads, issued together.
  const bool rec = MODE == 1 && !STATS && a.near != nullptr;
  float4 nd, ph = make_float4(0.f, 0.f, 0.f, 0.f);
  double nv0 = 0.0, nv1 = 0.0, nv2 = 0.0;
  uint64_t box = 0;
  if (rec) {
    const DevNearRec &r = a.near[cur];
    nd = r.node;
    nv0 = r.nv[0];
    nv1 = r.nv[1];
    nv2 = r.nv[2];
    box = r.nbox;
  } else {
    nd = reinterpret_cast<const float4 *>(a.nodes)[cur];
    ph = reinterpret_cast<const float4 *>(a.photons)[cur];
    if (MODE == 1 && !STATS && a.nbox) box = a.nbox[cur];
  }
  if (MODE == 1 && !STATS && (rec || a.nbox) && a.filter != 0 && normals_fail(box, q)) {
    // no photon of this subtree passes the PhotonFilter's normal test: none can enter the result
    // (the statistics instance walks it, counting the reference's visits)
    far_exists = false;
    return false;
  }
  double dx = (double)nd.x - q.q0, dy = (double)nd.y - q.q1, dz = (double)nd.z - q.q2;
  double d2 = dx * dx + dy * dy + dz * dz;
  if (STATS) ++cnt.visits;
  if (d2 < max_d2) {
    const uint32_t codes = __float_as_uint(ph.w);
    if (STATS) ++cnt.filtered;
    bool pass = a.filter == 0 || (rec ? photon_filter_nv(q, nv0, nv1, nv2, dx, dy, dz)
                                      : photon_filter(a, q, codes, dx, dy, dz));
    if (pass) {
      if (MODE == 1) {
        nearest = cur;
        max_d2 = d2;
      } else if (!STORE && size + 1 == k) {
        filled = true;
      } else if (size < k) {
        if (STORE) heap.set(size, cur, d2);
        ++size;
        if (SUMS) {
          if (d2 > sums.mx) sums.mx = d2;
          const double *dv = a.cdir + 3 * (size_t)(codes & 0xFFFFu);
          const bool ext = dv[0] * q.n0 + dv[1] * q.n1 + dv[2] * q.n2 > 0.0;
          add_split(ext, ph, sums.ext0, sums.ext1, sums.ext2, sums.in0, sums.in1, sums.in2);
        }
        if (STORE && size == k) {
          if (SUMS) sums.full = true;
          make_heap(heap, size);
          max_d2 = heap.d(0);
        }
      } else if (STORE) {
        pop_heap(heap, size);
        heap.set(k - 1, cur, d2);
        push_heap(heap, size);
        max_d2 = heap.d(0);
      }
    }
  }
  const uint32_t w = __float_as_uint(nd.w);
  axis = w & 3u;
  far_exists = false;
  if (axis == 3u) return false;  // leaf
  const double qa = axis_coord(q, axis);
  // a value select (a conditional over the members would select their addresses: scratch)
  const float sx = nd.x, sy = nd.y, sz = nd.z;
  split = axis == 0 ? (float)sx : (axis == 1 ? (float)sy : (float)sz);
  const bool has_left = (w & 4u) != 0;
  const uint32_t right = w >> 3;
  const bool has_right = right < a.n_nodes;
  const bool left_near = qa <= (double)split;
  const bool near_exists = left_near ? has_left : has_right;
  far_exists = left_near ? has_right : has_left;
  far = left_near ? right : cur + 1;
  if (near_exists) cur = left_near ? cur + 1 : right;
  return near_exists;
}

// The whole lookup of one query. The stack holds pending far sides as (far child << 2) | axis.
// A far side is pushed only when its plane is inside the radius at push time, and re-tested when
// popped only if the radius has shrunk since the lookup began (the plane coordinate then comes
// from psplit, the parent's split per node): the radius only shrinks, so a side that fails at
// push time fails at pop time too, and while the radius is unchanged the push-time test stands.
// This visits the nodes the recursion visits (KDTree.h:558-575), in its order, without
// re-reading the parent node at each pop.
// Without STORE the lookup returns true, with its state incomplete, at the insertion that would
// fill the set.
template <int MODE, bool SUMS, bool STORE, bool STATS, class Stack>
__device__ __forceinline__ bool lookup(const GatherArgs &a, const Query &q, double &max_d2, const Stack &stk,
                                       uint32_t &size, int64_t &nearest, const Heap &heap, DiscSums &sums,
                                       Counts &cnt) {
  if (a.n_nodes == 0) return false;
  const double r2 = max_d2;
  uint32_t sp = 0;
  uint32_t cur = 0;
  bool have = true;
  while (true) {
    if (!have) {
      bool next = false;
      while (sp > 0) {
        const uint32_t e = stk.get(--sp);
        const uint32_t far = e >> 2;
        if (max_d2 == r2) {
          cur = far;
          next = true;
          break;
        }
        const double qa = axis_coord(q, e & 3u);
        const double sd = (double)a.psplit[far];
        const double pd2 = (qa - sd) * (qa - sd);
        if (pd2 < max_d2) {
          cur = far;
          next = true;
          break;
        }
      }
      if (!next) break;
    }
    bool far_exists, filled = false;
    uint32_t far, axis;
    float split;
    have = visit<MODE, SUMS, STORE, STATS>(a, q, cur, max_d2, size, nearest, heap, sums, far_exists, far, axis,
                                          split, filled, cnt);
    if (!STORE && filled) return true;
    if (far_exists) {
      const double qa = axis_coord(q, axis);
      const double sd = (double)split;
      const double pd2 = (qa - sd) * (qa - sd);
      if (pd2 < max_d2) stk.set(sp++, (far << 2) | axis);
    }
  }
  return false;
}

// Cone-filtered irradiance over the set (Jensen's cone filter; not an estimator of the
// reference): photon weight 1 - d / (kc * r_max), only photons arriving from the normal's side,
// normalised by (1 - 2 / (3 kc)) * pi * paths * r_max^2, with r_max^2 = the set's largest d^2, or
// r^2 when the set is not full.
__device__ __forceinline__ void cone_estimate(const GatherArgs &a, const Query &q, uint64_t i, Heap heap, uint32_t size) {
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
__device__ __forceinline__ void caustic_estimate(const GatherArgs &a, const Query &q, uint64_t i, Heap heap, uint32_t size) {
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

__device__ __forceinline__ Query load_query(const GatherArgs &a, uint64_t i) {
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
  return q;
}

__device__ __forceinline__ Heap query_heap(const GatherArgs &a, uint64_t i) {
  return Heap{a.heap_pos + i * a.q_stride, a.heap_d2 + i * a.q_stride, a.j_stride};
}

// Outputs of query i once its lookup has ended: the nearest photon (MODE 1), or the set's size,
// ids, d^2 and the requested estimate (MODE 0).
// (Values, not references: a call that is not inlined would otherwise keep the caller's walk
// state in scratch memory.)
template <int MODE>
__device__ __forceinline__ void finish_query(const GatherArgs &a, const Query q, uint64_t i, const Heap heap,
                                             uint32_t size, int64_t nearest, const DiscSums sums) {
  if (MODE == 1) {
    a.nearest[i] = nearest < 0 ? kMiss : a.perm[nearest];
    return;
  }
  if (a.count) a.count[i] = size;
  if (a.ids || a.d2) {
    for (uint32_t j = 0; j < size; ++j) {
      uint32_t pos = heap.p(j);
      if (a.ids)
        a.ids[i * a.k + j] = a.ids_tree == 2   ? (a.photons[pos].codes & 0xFFFFu)
                             : a.ids_tree == 1 ? pos
                                               : a.perm[pos];
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
  double ext0 = sums.ext0, ext1 = sums.ext1, ext2 = sums.ext2, in0 = sums.in0, in1 = sums.in1, in2 = sums.in2;
  double mx = sums.mx;
  if (sums.full) ext0 = ext1 = ext2 = in0 = in1 = in2 = mx = 0.0;
  for (uint32_t j = 0; sums.full && j < size; ++j) {
    uint32_t pos = heap.p(j);
    double t = heap.d(j);
    if (t > mx) mx = t;
    float4 ph = reinterpret_cast<const float4 *>(a.photons)[pos];
    uint32_t dcode = __float_as_uint(ph.w) & 0xFFFFu;
    const double *dv = a.cdir + 3 * (size_t)dcode;
    double c = dv[0] * q.n0 + dv[1] * q.n1 + dv[2] * q.n2;
    add_split(c > 0.0, ph, ext0, ext1, ext2, in0, in1, in2);
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

// An upper bound for a nearest lookup (MODE 1) from the photons on the query's near path (the
// first nodes the lookup visits): the least d^2 below r2 among them that passes the filter. The
// nearest is the first photon, in the lookup's fixed visiting order, whose d^2 is the least of
// those below the start radius (a pruned side holds only photons at or beyond the radius, which
// cannot replace the current one under the strict comparison); every photon at or below the bound
// c is below the next double after c, so a lookup started there returns the same photon and d^2,
// and prunes far sides from its first node on. Returns r2 when no photon on the path qualifies.
__device__ __forceinline__ double near_path_bound(const GatherArgs &a, const Query &q, double r2) {
  double best = r2;
  bool any = false;
  uint32_t cur = 0;
  for (uint32_t depth = 0; depth < 64; ++depth) {
    float4 nd;
    bool from_rec = false;
    double nv0 = 0.0, nv1 = 0.0, nv2 = 0.0;
    uint32_t codes = 0;
    if (a.near) {
      const DevNearRec &r = a.near[cur];
      nd = r.node;
      nv0 = r.nv[0];
      nv1 = r.nv[1];
      nv2 = r.nv[2];
      from_rec = true;
    } else {
      nd = reinterpret_cast<const float4 *>(a.nodes)[cur];
      codes = __float_as_uint(reinterpret_cast<const float4 *>(a.photons)[cur].w);
    }
    const double dx = (double)nd.x - q.q0, dy = (double)nd.y - q.q1, dz = (double)nd.z - q.q2;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < best && (a.filter == 0 || (from_rec ? photon_filter_nv(q, nv0, nv1, nv2, dx, dy, dz)
                                                        : photon_filter(a, q, codes, dx, dy, dz)))) {
      best = d2;
      any = true;
    }
    const uint32_t w = __float_as_uint(nd.w);
    const uint32_t axis = w & 3u;
    if (axis == 3u) break;
    const float sx = nd.x, sy = nd.y, sz = nd.z;
    const float split = axis == 0 ? sx : (axis == 1 ? sy : sz);
    const bool left_near = axis_coord(q, axis) <= (double)split;
    const uint32_t right = w >> 3;
    const bool near_exists = left_near ? (w & 4u) != 0 : right < a.n_nodes;
    if (!near_exists) break;
    cur = left_near ? cur + 1 : right;
  }
  // the next double above best (best is a finite non-negative d^2 here)
  return any ? __longlong_as_double(__double_as_longlong(best) + 1) : r2;
}

// The sums are kept during the lookup only for the disc estimate.
__device__ __forceinline__ bool disc_sums(const GatherArgs &a) { return a.estimate == 0 && a.e_ext != nullptr; }

// One query from start to end (one lane, no refill).
// With a.defer (a disc estimate whose set is not returned) the lookup first runs without writing
// the set's slots: a set that never fills needs only the sums kept during insertion, which are the
// estimate's sums bit for bit. Only a lookup whose set fills is run again from the start with its
// slots written (the same visits in the same order, then make_heap and the bounded heap), so the
// slots cost memory traffic only for sets that use them.
// STATS writes the lookup's work counts (visits, photons inside the radius, final set size) to
// a.stats[3 i ..]; that instance always writes the slots.
template <int MODE, bool STATS, class Stack>
__device__ void gather_one(const GatherArgs &a, uint64_t t, const Stack &stk) {
  const uint64_t i = a.order ? (uint64_t)a.order[t] : t;
  const Query q = load_query(a, i);
  double max_d2 = a.r * a.r;
  uint32_t size = 0;
  int64_t nearest = -1;
  const Heap heap = query_heap(a, i);
  DiscSums sums{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false};
  Counts cnt{0u, 0u};
  if (MODE == 1) {
    // (the statistics instance counts the visits of the lookup from the full radius)
    if (!STATS && a.n_nodes > 0) max_d2 = near_path_bound(a, q, max_d2);
    lookup<1, false, true, STATS>(a, q, max_d2, stk, size, nearest, heap, sums, cnt);
  } else if (a.k > 0) {
    if (disc_sums(a)) {
      bool filled = false;
      if (!STATS && a.defer) filled = lookup<0, true, false, false>(a, q, max_d2, stk, size, nearest, heap, sums, cnt);
      if (STATS || !a.defer || filled) {
        max_d2 = a.r * a.r;
        size = 0;
        sums = DiscSums{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false};
        lookup<0, true, true, STATS>(a, q, max_d2, stk, size, nearest, heap, sums, cnt);
      }
    } else {
      lookup<0, false, true, STATS>(a, q, max_d2, stk, size, nearest, heap, sums, cnt);
    }
  }
  if (STATS) {
    a.stats[3 * i] = cnt.visits;
    a.stats[3 * i + 1] = cnt.filtered;
    a.stats[3 * i + 2] = MODE == 1 ? (nearest < 0 ? 0u : 1u) : size;
  }
  finish_query<MODE>(a, q, i, heap, size, nearest, sums);
}

}  // namespace
}  // namespace gath
}  // namespace skw
