// Device-side traversal of the ternary triangle tree, shared by the single-scene kernels
// (traverse.hip) and the multi-scene batch kernel (batch.hip).
//
// Semantics follow TriangleAccelerator::_NodeIntersect / _NodeIntersectTest
// (Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:154-293, 354-462) bit for bit:
//  * FP64 slab test of the float node box with the 12 comparisons of :188-192 (a NaN never culls);
//  * children visited left, middle, right when dir[axis] > 0.0, else right, middle, left (:199-210),
//    each tested against the ray's current tmax when it is reached;
//  * FP64 Moller-Trumbore on float vertices with the +-DBL_EPSILON bounds of :256/:261; an accepted
//    t (tmin <= t <= tmax) becomes the new tmax, so on an exact tie the later triangle wins (:265);
//  * instances of a leaf are traversed before its triangles with the ray moved by the instance's
//    world-to-mesh matrix (:215-233).
// Every translation unit including this is compiled with -ffp-contract=off (no fused multiply-add,
// as in the reference's SSE2 build); f64 division stays the correctly rounded IEEE sequence.
//
// Layout: nodes are 32 bytes, children of a node form one contiguous block, so a stack entry is
// one 32-bit word: next sibling index (29 bits), walk direction (1 bit), siblings left (2 bits).
// The stack lives in LDS, [depth][lane], so every lane hits its own bank at any depth.
#pragma once
#include <hip/hip_runtime.h>

#include "../skw_layout.hpp"
#include "../../../include/skwarka_rt.h"
#include "kernels.hpp"

namespace skw {
namespace trace {


constexpr double kEps = 2.220446049250313080847e-16;  // DBL_EPSILON (Constants.h: numeric_limits epsilon)
constexpr int kWave = 64;

struct RayState {
  double o0, o1, o2, d0, d1, d2, tmin;
  double i0, i1, i2;
};

struct Best {
  uint32_t pos;   // tree position of the best plain triangle
  uint32_t prim;  // for instance hits
  uint32_t tri;
  bool inst;
  bool hit;
  double b1, b2;
};

// Bit k set when direction component k is > 0.0: the child order rule of :199 as integer bits,
// so the split axis selects a bit instead of a double (a selected double would push the ray into
// scratch memory).
__device__ __forceinline__ uint32_t dir_mask(const RayState &r) {
  return (r.d0 > 0.0 ? 1u : 0u) | (r.d1 > 0.0 ? 2u : 0u) | (r.d2 > 0.0 ? 4u : 0u);
}

__device__ __forceinline__ void load_node(const DevNode *__restrict__ nodes, uint32_t idx, float4 &lo, float4 &hi) {
  const float4 *p = reinterpret_cast<const float4 *>(nodes + idx);
  lo = p[0];
  hi = p[1];
}

// Makes a loaded node word live at this point, so the compiler issues the whole 32-byte node load
// at once instead of sinking the child words' load into the branch that reads them (which would
// add a second dependent memory round trip to the step). Emits no instruction.
__device__ __forceinline__ void keep_live(const float4 &v) { asm volatile("" ::"v"(v.z), "v"(v.w)); }

// true = culled. TriangleAccelerator.cpp:176-192.
__device__ __forceinline__ bool slab_cull(const float4 &lo, const float4 &hi, const RayState &r, double tmax) {
  double tn1 = ((double)lo.x - r.o0) * r.i0;
  double tf1 = ((double)lo.w - r.o0) * r.i0;
  bool s1 = tn1 > tf1;
  double a1 = s1 ? tf1 : tn1, b1 = s1 ? tn1 : tf1;
  double tn2 = ((double)lo.y - r.o1) * r.i1;
  double tf2 = ((double)hi.x - r.o1) * r.i1;
  bool s2 = tn2 > tf2;
  double a2 = s2 ? tf2 : tn2, b2 = s2 ? tn2 : tf2;
  double tn3 = ((double)lo.z - r.o2) * r.i2;
  double tf3 = ((double)hi.y - r.o2) * r.i2;
  bool s3 = tn3 > tf3;
  double a3 = s3 ? tf3 : tn3, b3 = s3 ? tn3 : tf3;
  return r.tmin > b1 || a1 > b2 || a2 > b1 || a3 > b1 || r.tmin > b2 || a1 > b3 || a2 > b3 || a3 > b2 ||
         r.tmin > b3 || a1 > tmax || a2 > tmax || a3 > tmax;
}

// A triangle's nine float vertex coordinates (the first 36 bytes of its 48-byte record).
struct TriVerts {
  float4 q0, q1;
  float q2x;
};

__device__ __forceinline__ TriVerts load_tri(const DevTri *__restrict__ tris, uint32_t pos) {
  const float4 *p = reinterpret_cast<const float4 *>(tris + pos);
  TriVerts v;
  v.q0 = p[0];
  v.q1 = p[1];
  v.q2x = reinterpret_cast<const float *>(p + 2)[0];
  return v;
}

// The same decision as slab_cull for a ray whose slab products cannot be NaN. Box coordinates are
// never NaN (they may be +-inf), so a product (coord - o) * inv is NaN only for 0 * inf: a zero or
// tiny direction component with the origin on the slab plane, an infinite direction component
// (inv = 0) against an infinite coordinate, or a non-finite origin. For a ray with finite origin,
// direction and inverse direction every product is a number (or +-inf), so the swap of :183-185
// is min / max and the 12
// comparisons of :188-192 are max(tmin, a1, a2, a3) > min(b1, b2, b3) or max(a1, a2, a3) > tmax
// (the pairs a_k > b_k they add are false since a_k <= b_k, and tmin > tmax is not one of them).
__device__ __forceinline__ bool slab_cull_nan_free(const float4 &lo, const float4 &hi, const RayState &r,
                                                   double tmax) {
  const double tn1 = ((double)lo.x - r.o0) * r.i0, tf1 = ((double)lo.w - r.o0) * r.i0;
  const double tn2 = ((double)lo.y - r.o1) * r.i1, tf2 = ((double)hi.x - r.o1) * r.i1;
  const double tn3 = ((double)lo.z - r.o2) * r.i2, tf3 = ((double)hi.y - r.o2) * r.i2;
  const double amax = fmax(fmax(fmin(tn1, tf1), fmin(tn2, tf2)), fmin(tn3, tf3));
  const double bmin = fmin(fmin(fmax(tn1, tf1), fmax(tn2, tf2)), fmax(tn3, tf3));
  return fmax(amax, r.tmin) > bmin || amax > tmax;
}

__device__ __forceinline__ bool ray_nan_free(const RayState &r) {
  return isfinite(r.o0) && isfinite(r.o1) && isfinite(r.o2) && isfinite(r.d0) && isfinite(r.d1) &&
         isfinite(r.d2) && isfinite(r.i0) && isfinite(r.i1) && isfinite(r.i2);
}

// The child words (first child index, kind / mask / count) of a node: its last 8 bytes.
__device__ __forceinline__ uint2 load_child_words(const DevNode *__restrict__ nodes, uint32_t idx) {
  return reinterpret_cast<const uint2 *>(nodes + idx)[3];
}

// TriangleAccelerator.cpp:238-274. Returns true with t, b1, b2 when accepted in [tmin, tmax].
__device__ __forceinline__ bool tri_test_verts(const TriVerts &tv, const RayState &r, double tmax, double &t_out,
                                               double &b1_out, double &b2_out) {
  const float4 q0 = tv.q0, q1 = tv.q1;
  const float4 q2 = make_float4(tv.q2x, 0.f, 0.f, 0.f);
  double v0x = q0.x, v0y = q0.y, v0z = q0.z;
  double e1x = (double)q0.w - v0x, e1y = (double)q1.x - v0y, e1z = (double)q1.y - v0z;
  double e2x = (double)q1.z - v0x, e2y = (double)q1.w - v0y, e2z = (double)q2.x - v0z;
  // s1 = dir ^ e2
  double s1x = r.d1 * e2z - r.d2 * e2y;
  double s1y = r.d2 * e2x - r.d0 * e2z;
  double s1z = r.d0 * e2y - r.d1 * e2x;
  double divisor = s1x * e1x + s1y * e1y + s1z * e1z;
  double inv = 1.0 / divisor;
  double dx = r.o0 - v0x, dy = r.o1 - v0y, dz = r.o2 - v0z;
  double b1 = (dx * s1x + dy * s1y + dz * s1z) * inv;
  if (b1 > -kEps && b1 < (1.0 + kEps)) {
    // s2 = d ^ e1
    double s2x = dy * e1z - dz * e1y;
    double s2y = dz * e1x - dx * e1z;
    double s2z = dx * e1y - dy * e1x;
    double b2 = (r.d0 * s2x + r.d1 * s2y + r.d2 * s2z) * inv;
    if (b2 > -kEps && b1 + b2 < (1.0 + kEps)) {
      double t = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
      if (t >= r.tmin && t <= tmax) {
        t_out = t;
        b1_out = b1;
        b2_out = b2;
        return true;
      }
    }
  }
  return false;
}

__device__ __forceinline__ bool tri_test(const DevTri *__restrict__ tris, uint32_t pos, const RayState &r,
                                         double tmax, double &t_out, double &b1_out, double &b2_out) {
  return tri_test_verts(load_tri(tris, pos), r, tmax, t_out, b1_out, b2_out);
}

__device__ __forceinline__ void make_state(RayState &s, double o0, double o1, double o2, double d0, double d1,
                                           double d2, double tmin) {
  s.o0 = o0;
  s.o1 = o1;
  s.o2 = o2;
  s.d0 = d0;
  s.d1 = d1;
  s.d2 = d2;
  s.tmin = tmin;
  s.i0 = 1.0 / d0;
  s.i1 = 1.0 / d1;
  s.i2 = 1.0 / d2;
}

// Ray moved into an instance's mesh space (Transform.h:223-300): origin as a point (with the w
// division when w != 1), direction as a vector; tmin/tmax unchanged.
__device__ __forceinline__ void instance_ray(const DevInst *__restrict__ inst, const RayState &r, RayState &t) {
  const double *m = inst->w2m;
  double xp = m[0] * r.o0 + m[1] * r.o1 + m[2] * r.o2 + m[3];
  double yp = m[4] * r.o0 + m[5] * r.o1 + m[6] * r.o2 + m[7];
  double zp = m[8] * r.o0 + m[9] * r.o1 + m[10] * r.o2 + m[11];
  double w = m[12] * r.o0 + m[13] * r.o1 + m[14] * r.o2 + m[15];
  if (w != 1.0) {
    double iw = 1.0 / w;
    xp = xp * iw;
    yp = yp * iw;
    zp = zp * iw;
  }
  double dx = m[0] * r.d0 + m[1] * r.d1 + m[2] * r.d2;
  double dy = m[4] * r.d0 + m[5] * r.d1 + m[6] * r.d2;
  double dz = m[8] * r.d0 + m[9] * r.d1 + m[10] * r.d2;
  make_state(t, xp, yp, zp, dx, dy, dz, r.tmin);
}

// Stack of 32-bit entries for one lane. LDS variant: [depth][lane] inside the workgroup's array.
template <int S>
struct LdsStack {
  uint32_t *base;  // &lds[lane]
  __device__ __forceinline__ uint32_t get(uint32_t i) const { return base[i * kWave]; }
  __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { base[i * kWave] = v; }
};
// Global-memory variant for trees deeper than the LDS stack: [depth][global lane].
struct GlobalStack {
  uint32_t *base;
  uint64_t stride;
  __device__ __forceinline__ uint32_t get(uint32_t i) const { return base[i * stride]; }
  __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { base[i * stride] = v; }
};

// A complete walk of an instanced mesh's subtree (triangles only, one level of instancing as in
// TriangleAccelerator.cpp:215-233) with the stack above `sp0`. ANY: stop at the first accepted
// triangle. Updates tmax and best.pos / b1 / b2 on each accepted triangle.
template <bool ANY, bool STATS, class Stack>
__device__ bool walk_instance(const DevNode *__restrict__ nodes, const DevTri *__restrict__ tris, uint32_t start,
                              const RayState &r, double &tmax, const Stack &stk, uint32_t sp0, Best &best,
                              uint32_t &pops, uint32_t &tests) {
  uint32_t sp = sp0;
  uint32_t cur = start;
  bool have = true;
  bool found = false;
  const uint32_t dpos = dir_mask(r);
  while (true) {
    if (!have) {
      if (sp == sp0) break;
      uint32_t e = stk.get(sp - 1);
      cur = e >> 3;
      uint32_t left = e & 3u;
      if (left > 1) {
        uint32_t nxt = (e & 4u) ? cur - 1 : cur + 1;
        stk.set(sp - 1, (nxt << 3) | (e & 4u) | (left - 1));
      } else {
        --sp;
      }
    }
    have = false;
    float4 lo, hi;
    load_node(nodes, cur, lo, hi);
    if (STATS) ++pops;
    if (slab_cull(lo, hi, r, tmax)) continue;
    uint32_t a = __float_as_uint(hi.z), b = __float_as_uint(hi.w);
    uint32_t kind = b & 3u;
    if (kind != 0) {
      uint32_t nch = __builtin_popcount((b >> 2) & 7u);
      if ((dpos >> (kind - 1u)) & 1u) {
        cur = a;
        if (nch > 1) stk.set(sp++, ((a + 1) << 3) | (nch - 1));
      } else {
        cur = a + nch - 1;
        if (nch > 1) stk.set(sp++, ((cur - 1) << 3) | 4u | (nch - 1));
      }
      have = true;
      continue;
    }
    uint32_t cnt = b >> 3;
    for (uint32_t k = 0; k < cnt; ++k) {
      double t, b1, b2;
      if (STATS) ++tests;
      if (tri_test(tris, a + k, r, tmax, t, b1, b2)) {
        if (ANY) return true;
        tmax = t;
        found = true;
        best.pos = a + k;
        best.b1 = b1;
        best.b2 = b2;
      }
    }
  }
  return found;
}

__device__ __forceinline__ void load_ray(const skw_ray *__restrict__ rays, uint64_t i, RayState &s, double &tmax) {
  const double2 *p = reinterpret_cast<const double2 *>(rays + i);
  double2 a = p[0], b = p[1], c = p[2], d = p[3];
  make_state(s, a.x, a.y, b.x, b.y, c.x, c.y, d.x);
  tmax = d.y;
}

__device__ __forceinline__ bool zero_signed_or_nonfinite(double x) {
  return (unsigned long long)__double_as_longlong(x) == 0x8000000000000000ull || !isfinite(x);
}

// Intersect() recomputes b1/b2 after moving the ray to mesh space with the primitive's
// world-to-mesh matrix (TriangleAccelerator.cpp:311-322). For the identity matrix that product,
// (1*x + 0*y) + 0*z (+ 0 for points), returns x unchanged except that -0.0 becomes +0.0 and a
// non-finite component in the row poisons it. Rays with such components get their barycentrics
// recomputed here with exactly that arithmetic; all other rays keep the traversal values, which
// are then identical.
__device__ __forceinline__ void identity_bary(const DevTri *__restrict__ tris, uint32_t pos, const RayState &r,
                                              double &b1, double &b2) {
  const double one = 1.0, zero = 0.0;
  double ox = one * r.o0 + zero * r.o1 + zero * r.o2 + zero;
  double oy = zero * r.o0 + one * r.o1 + zero * r.o2 + zero;
  double oz = zero * r.o0 + zero * r.o1 + one * r.o2 + zero;
  double w = zero * r.o0 + zero * r.o1 + zero * r.o2 + one;
  if (w != 1.0) {
    double iw = 1.0 / w;
    ox = ox * iw;
    oy = oy * iw;
    oz = oz * iw;
  }
  double dx = one * r.d0 + zero * r.d1 + zero * r.d2;
  double dy = zero * r.d0 + one * r.d1 + zero * r.d2;
  double dz = zero * r.d0 + zero * r.d1 + one * r.d2;
  const float4 *p = reinterpret_cast<const float4 *>(tris + pos);
  float4 q0 = p[0], q1 = p[1], q2 = p[2];
  double v0x = q0.x, v0y = q0.y, v0z = q0.z;
  double e1x = (double)q0.w - v0x, e1y = (double)q1.x - v0y, e1z = (double)q1.y - v0z;
  double e2x = (double)q1.z - v0x, e2y = (double)q1.w - v0y, e2z = (double)q2.x - v0z;
  double s1x = dy * e2z - dz * e2y;
  double s1y = dz * e2x - dx * e2z;
  double s1z = dx * e2y - dy * e2x;
  double divisor = s1x * e1x + s1y * e1y + s1z * e1z;
  b1 = 0.0;
  b2 = 0.0;
  if (divisor == 0.0) return;  // TriangleMesh.cpp:421
  double inv = 1.0 / divisor;
  double ddx = ox - v0x, ddy = oy - v0y, ddz = oz - v0z;
  b1 = (ddx * s1x + ddy * s1y + ddz * s1z) * inv;
  double s2x = ddy * e1z - ddz * e1y;
  double s2y = ddz * e1x - ddx * e1z;
  double s2z = ddx * e1y - ddy * e1x;
  b2 = (dx * s2x + dy * s2y + dz * s2z) * inv;
}


// Closest hit record of a finished walk (skw_hit layout, two 16-byte stores).
__device__ __forceinline__ void store_hit(skw_hit *out, const DevTri *__restrict__ tris, bool hit, const Best &best,
                                          double tmax, const RayState &r) {
  uint32_t prim = kMiss, tri = kMiss;
  double t = __builtin_huge_val(), b1 = 0.0, b2 = 0.0;
  if (hit) {
    if (best.inst) {
      prim = best.prim;
      tri = best.tri;
    } else {
      const uint4 ids = reinterpret_cast<const uint4 *>(tris + best.pos)[2];
      prim = ids.y;
      tri = ids.z;
    }
    t = tmax;
    b1 = best.b1;
    b2 = best.b2;
    if (!best.inst && (zero_signed_or_nonfinite(r.o0) || zero_signed_or_nonfinite(r.o1) ||
                       zero_signed_or_nonfinite(r.o2) || zero_signed_or_nonfinite(r.d0) ||
                       zero_signed_or_nonfinite(r.d1) || zero_signed_or_nonfinite(r.d2)))
      identity_bary(tris, best.pos, r, b1, b2);
  }
  uint4 w0;
  w0.x = prim;
  w0.y = tri;
  unsigned long long tb = (unsigned long long)__double_as_longlong(t);
  w0.z = (uint32_t)tb;
  w0.w = (uint32_t)(tb >> 32);
  uint4 *o = reinterpret_cast<uint4 *>(out);
  o[0] = w0;
  reinterpret_cast<double2 *>(o)[1] = make_double2(b1, b2);
}

// One lane's walk as a resumable state machine. start() tests the root. advance() makes one move
// between leaves: with no current node it pops the next pending sibling and re-tests its box; an
// internal current node has all its children tested at once with the current tmax, the first
// survivor is entered and later survivors pushed as one entry. When the current node is an
// accepted leaf, advance() stops there (at_leaf) until leaf() has tested its instances and
// triangles; the caller chooses when, so that lanes of a wave test leaves together. advance()
// returns true when the walk has ended; found / best / tmax then hold the result.
// Testing the children at visit time is exact: the first survivor is the node the reference pops
// next, with the same tmax; a culled child stays culled at its later pop, since its tmax-free
// conditions do not change and tmax only shrinks; later survivors are tested again when popped.
// Holding a lane at a leaf changes no value: each lane's own sequence of tests is unchanged.
template <bool ANY, bool INST, bool STATS, class Stack>
struct Walker {
  const DevNode *__restrict__ nodes;
  const DevTri *__restrict__ tris;
  const DevInst *__restrict__ insts;
  const uint32_t *__restrict__ inst_range;
  uint32_t root;
  float4 root_lo, root_hi;  // the root's box and child words, loaded once per tree by set_root()
  Stack stk;
  RayState r;
  double tmax;
  Best best;
  uint32_t sp, cur, ca, cb, dpos;
  bool have, found, at_leaf;
  bool nan_free;  // slab products of this ray can never be NaN (see slab_cull_nan_free)
  uint32_t pops, tests;

  __device__ __forceinline__ void set_root(const DevNode *__restrict__ tree_nodes, uint32_t root_idx) {
    nodes = tree_nodes;
    root = root_idx;
    load_node(tree_nodes, root_idx, root_lo, root_hi);
  }

  __device__ __forceinline__ void start(const RayState &ray, double tmax0) {
    r = ray;
    tmax = tmax0;
    dpos = dir_mask(ray);
    nan_free = ray_nan_free(ray);
    best.hit = false;
    best.inst = false;
    sp = 0;
    found = false;
    pops = 0;
    tests = 0;
    // The root is tested like any popped node (TriangleAccelerator.cpp:165-192).
    if (STATS) ++pops;
    have = !slab_cull(root_lo, root_hi, r, tmax);
    cur = root;
    ca = __float_as_uint(root_hi.z);
    cb = __float_as_uint(root_hi.w);
    at_leaf = have && (cb & 3u) == 0;
  }

  // FAST: every busy lane of the wave has nan_free set and the tree's boxes nest (a wave-uniform
  // choice of the slab test and of the pop below).
  template <bool FAST>
  __device__ __forceinline__ bool cull(const float4 &lo, const float4 &hi) const {
    return FAST ? slab_cull_nan_free(lo, hi, r, tmax) : slab_cull(lo, hi, r, tmax);
  }

  template <bool FAST>
  __device__ __forceinline__ bool advance() {
    // A popped leaf in the FAST closest-hit walk: its box is the only one tested below.
    bool single = false;
    if (!have) {
      if (sp == 0) return true;
      uint32_t e = stk.get(sp - 1);
      cur = e >> 3;
      uint32_t left = e & 3u;
      if (left > 1) {
        uint32_t nxt = (e & 4u) ? cur - 1 : cur + 1;
        stk.set(sp - 1, (nxt << 3) | (e & 4u) | (left - 1));
      } else {
        --sp;
      }
      if (FAST && !ANY && !STATS) {
        // The popped node's own re-test is left out when it has children. Its children's boxes
        // lie inside its box, and for a NaN-free ray every slab value is monotone in the box
        // coordinates (rounding is monotone), so each child's entry and exit values lie between
        // the parent's: a parent culled with the current tmax has every child culled with it.
        // Testing the children directly (as below) therefore reaches the same nodes as testing
        // the parent first. The counting walk (STATS) keeps the reference's test of every node.
        const uint2 cw = load_child_words(nodes, cur);
        ca = cw.x;
        cb = cw.y;
        single = (cb & 3u) == 0;
      } else {
        float4 lo, hi;
        load_node(nodes, cur, lo, hi);
        keep_live(hi);
        // An occlusion walk never changes tmax (it ends at its first hit), so a sibling that
        // passed its test when pushed passes it again now: only the closest-hit walk re-tests.
        if (!ANY && cull<FAST>(lo, hi)) return false;
        ca = __float_as_uint(hi.z);
        cb = __float_as_uint(hi.w);
        if ((cb & 3u) == 0) {
          have = true;
          at_leaf = true;
          return false;
        }
      }
    }
    have = false;
    const uint32_t a = ca, b = cb;
    const uint32_t kind = single ? 1u : b & 3u;
    const uint32_t nch = single ? 1u : __builtin_popcount((b >> 2) & 7u);
    const bool fwd = single || ((dpos >> (kind - 1u)) & 1u);
    // Children in visiting order (scalars, not arrays: an indexed array would live in scratch).
    // Absent children alias c0, so the three loads are issued together and wait once.
    const uint32_t c0 = single ? cur : (fwd ? a : a + nch - 1u);
    const uint32_t c1 = nch > 1 ? (fwd ? a + 1u : a + nch - 2u) : c0;
    const uint32_t c2 = nch > 2 ? a + (fwd ? 2u : 0u) : c0;
    float4 lo0, hi0, lo1, hi1, lo2, hi2;
    load_node(nodes, c0, lo0, hi0);
    load_node(nodes, c1, lo1, hi1);
    load_node(nodes, c2, lo2, hi2);
    keep_live(hi0);
    keep_live(hi1);
    keep_live(hi2);
    // All three tests are evaluated (no short circuit), so no load is sunk into a branch.
    const uint32_t k0 = cull<FAST>(lo0, hi0) ? 0u : 1u;
    const uint32_t k1 = cull<FAST>(lo1, hi1) ? 0u : 2u;
    const uint32_t k2 = cull<FAST>(lo2, hi2) ? 0u : 4u;
    const uint32_t surv = (k0 | k1 | k2) & ((1u << nch) - 1u);
    if (STATS) pops += nch;
    if (surv == 0) return false;
    const uint32_t s0 = __builtin_ctz(surv);
    cur = s0 == 0 ? c0 : (s0 == 1 ? c1 : c2);
    const float hz = s0 == 0 ? hi0.z : (s0 == 1 ? hi1.z : hi2.z);
    const float hw = s0 == 0 ? hi0.w : (s0 == 1 ? hi1.w : hi2.w);
    ca = __float_as_uint(hz);
    cb = __float_as_uint(hw);
    have = true;
    at_leaf = (cb & 3u) == 0;
    const uint32_t rest = surv & ~((2u << s0) - 1u);
    if (rest) {
      const uint32_t p = __builtin_ctz(rest);
      const uint32_t start = p == 1 ? c1 : c2;
      stk.set(sp++, (start << 3) | (fwd ? 0u : 4u) | (uint32_t)__builtin_popcount(rest));
    }
    return false;
  }

  // Tests the instances and triangles of the accepted leaf the lane is holding.
  __device__ __forceinline__ void leaf() {
    at_leaf = false;
    have = false;
    const uint32_t a = ca, b = cb;
    bool stop = false;
    if (INST && (b & 4u)) {
      uint32_t ib = inst_range[2 * cur], ie = inst_range[2 * cur + 1];
      for (uint32_t k = ib; k < ie && !stop; ++k) {
        const DevInst *in = insts + k;
        RayState tr;
        instance_ray(in, r, tr);
        Best nb;
        if (walk_instance<ANY, STATS>(nodes, tris, in->root, tr, tmax, stk, sp, nb, pops, tests)) {
          found = true;
          if (ANY) {
            stop = true;
          } else {
            best.hit = true;
            best.inst = true;
            best.prim = in->prim;
            best.tri = reinterpret_cast<const uint4 *>(tris + nb.pos)[2].z;
            best.b1 = nb.b1;
            best.b2 = nb.b2;
          }
        }
      }
    }
    // The leaf's triangles are contiguous: the next one's vertices are loaded while the current
    // one is tested (the last load re-reads the current triangle instead of running past the leaf).
    const uint32_t cnt = b >> 3;
    TriVerts nv;
    if (cnt) nv = load_tri(tris, a);  // an empty leaf's triangle index may be past the array
    for (uint32_t k = 0; k < cnt && !stop; ++k) {
      double t, b1, b2;
      if (STATS) ++tests;
      const TriVerts tv = nv;
      nv = load_tri(tris, a + (k + 1 < cnt ? k + 1 : k));
      if (tri_test_verts(tv, r, tmax, t, b1, b2)) {
        found = true;
        if (ANY) {
          stop = true;
        } else {
          tmax = t;
          best.hit = true;
          best.inst = false;
          best.pos = a + k;
          best.b1 = b1;
          best.b2 = b2;
        }
      }
    }
    if (ANY && stop) sp = 0;  // the walk ends at its first accepted hit
  }

  // The held leaves of the wave tested together; every lane of the wave calls it (hold: this lane
  // holds an accepted leaf). Leaves with instances or with more than kPoolMax triangles are tested
  // by their own lane as in leaf(). The other leaves' triangles are spread over the wave's lanes,
  // 64 per round: each is tested with its owner's ray and the owner's tmax at the round's start,
  // then every owner takes its own hits in its leaf's order and accepts t <= its current tmax.
  // That accepts exactly the triangles leaf() accepts: a test depends on tmax only through
  // t <= tmax, and tmax only shrinks along the leaf (so a rejected t stays rejected).
  // slots: kWave * kPoolMax bytes of LDS for this wave; slot j of a phase holds the owner lane of
  // the phase's j-th pooled triangle and the triangle's place in the owner's leaf.
  static constexpr uint32_t kPoolMax = 4;
  __device__ __forceinline__ void leaves_pooled(bool hold, uint8_t *slots) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t n_all = hold ? (cb >> 3) : 0u;
    const bool solo = hold && ((INST && (cb & 4u)) || n_all > kPoolMax);
    if (solo) leaf();
    const bool pooled = hold && !solo;
    const uint32_t cnt = pooled ? n_all : 0u;  // 0..kPoolMax: three bits
    // Exclusive prefix of the counts over the wave, one ballot per bit of the count.
    const uint64_t m1 = __ballot(cnt & 1u), m2 = __ballot(cnt & 2u), m4 = __ballot(cnt & 4u);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t off = (uint32_t)(__popcll(m1 & below) + 2 * __popcll(m2 & below) + 4 * __popcll(m4 & below));
    const uint32_t total = (uint32_t)(__popcll(m1) + 2 * __popcll(m2) + 4 * __popcll(m4));
    if (pooled) {
      at_leaf = false;
      have = false;
    }
    for (uint32_t k = 0; k < kPoolMax; ++k)
      if (k < cnt) slots[off + k] = (uint8_t)(lane | (k << 6));
    // The slots are read by other lanes of this wave: keep the writes before the reads.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bool stop = false;
    for (uint32_t base = 0; base < total; base += kWave) {
      const uint32_t j = base + lane;
      const uint32_t slot = j < total ? slots[j] : 0u;
      const uint32_t o = slot & (kWave - 1);
      RayState rr;
      rr.o0 = __shfl(r.o0, o);
      rr.o1 = __shfl(r.o1, o);
      rr.o2 = __shfl(r.o2, o);
      rr.d0 = __shfl(r.d0, o);
      rr.d1 = __shfl(r.d1, o);
      rr.d2 = __shfl(r.d2, o);
      rr.tmin = __shfl(r.tmin, o);
      const double otmax = __shfl(tmax, o);
      const uint32_t opos = __shfl(ca, o) + (slot >> 6);
      double t = 0.0, b1 = 0.0, b2 = 0.0;
      bool h = false;
      if (j < total) h = tri_test(tris, opos, rr, otmax, t, b1, b2);
      // This lane's pooled triangles in the round: bits [lo, hi) of the round's lanes.
      const int32_t s = (int32_t)off - (int32_t)base, e = s + (int32_t)cnt;
      const int32_t lo = s < 0 ? 0 : s, hi = e > (int32_t)kWave ? (int32_t)kWave : e;
      const uint64_t seg = hi > lo ? ((1ull << (hi - lo)) - 1ull) << lo : 0ull;
      const uint64_t hits = __ballot(h);  // every lane votes: a stopped walk still tests for others
      uint64_t mine = stop ? 0ull : hits & seg;
      if (ANY) {
        // the walk ends at its first accepted triangle
        if (STATS) tests += (uint32_t)__popcll(mine ? seg & ((mine & (0ull - mine)) * 2ull - 1ull) : (stop ? 0ull : seg));
        if (mine) {
          found = true;
          stop = true;
        }
      } else {
        if (STATS) tests += (uint32_t)__popcll(seg);
        // in leaf order, each accepted t becomes the bound for the next
        while (__ballot(mine != 0ull)) {
          const uint32_t src = mine ? (uint32_t)__ffsll((unsigned long long)mine) - 1u : lane;
          const double tt = __shfl(t, src), bb1 = __shfl(b1, src), bb2 = __shfl(b2, src);
          if (mine) {
            if (tt <= tmax) {
              found = true;
              tmax = tt;
              best.hit = true;
              best.inst = false;
              best.pos = ca + (base + src - off);
              best.b1 = bb1;
              best.b2 = bb2;
            }
            mine &= mine - 1ull;
          }
        }
      }
    }
    if (ANY && stop) sp = 0;  // the walk ends at its first accepted hit
  }

  // One move of either kind (for callers that do not group leaves); true when the walk has ended.
  __device__ __forceinline__ bool step() {
    if (at_leaf) {
      leaf();
      return false;
    }
    return advance<false>();
  }
};

}  // namespace trace
}  // namespace skw
