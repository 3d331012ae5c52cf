// ORACLE — TEST INFRASTRUCTURE ONLY (see orc_math.hpp). Never linked into the product library.
//
// CPU restatement of Skwarka's TriangleAccelerator (ternary kd-like tree over scene triangles)
// and of the hit attributes the reference derives from it:
//   constructor, mesh grouping, instancing   Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:22-147
//   _NodeIntersect (closest hit)                                   TriangleAccelerator.cpp:154-293
//   Intersect (+ differential geometry, m_cross, m_dot)            TriangleAccelerator.cpp:295-352
//   _NodeIntersectTest / IntersectTest (occlusion)                 TriangleAccelerator.cpp:354-468
//   _DetermineBestSplit (200 binned split candidates, SAH-like)    TriangleAccelerator.cpp:505-617
//   Node::Node (3-way in-place partition, middle-split mask)       TriangleAccelerator.cpp:621-729
//   constants MAX_TRIANGLES_IN_LEAF=4, MAX_SPLIT_TRIES=200, MAX_TREE_DEPTH=200  TriangleAccelerator.h:127-134
//   mesh cleaning (drops out-of-range / repeated-index triangles)  Source/RayLibs/Raytracer/Core/TriangleMesh.cpp:134-147,180-186
//   _ComputeIntersectionPoint / ComputeDifferentialGeometry        TriangleMesh.cpp:414-561
//   CoreUtils::GetNextMinT                                         Source/RayLibs/Raytracer/Core/CoreUtils.h:139-149
//   brute-force reference used by the reference's own tests       UnitTests/MainTests/Raytracer/Core/TriangleAccelerator.test.h:213-269
//
// Two choices are pinned where the reference leaves behaviour to the platform:
//  * mesh order: the reference walks a std::map keyed by mesh address; meshes here are walked in
//    the caller's mesh order (what an allocator that hands out rising addresses produces);
//  * (size_t) of a NaN bin position (a node with zero extent on an axis): the x86-64 SSE2
//    conversion yields 2^63, which the std::min clamps to num_tries.
#include <algorithm>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "../include/orc_api.h"
#include "orc_math.hpp"

namespace orc {

static const size_t MAX_TRIANGLES_IN_LEAF = 4;
static const size_t MAX_SPLIT_TRIES = 200;
static const size_t MAX_TREE_DEPTH = 200;
static const size_t NO_PRIM = (size_t)-1;

static inline size_t bin_of(double v, size_t num_tries) {
  if (!(v < 18446744073709551616.0)) return num_tries;  // NaN / +inf / too large: see header
  return smin(num_tries, (size_t)v);
}

struct Mesh {
  std::vector<F3> verts;
  std::vector<uint32_t> tris;     // 3 per triangle, after cleaning
  std::vector<F3> normals;        // optional per-vertex shading normals (normalised)
  bool invert_normals = false;
};

struct Node {
  BBoxF bbox;
  Node *ch[3] = {nullptr, nullptr, nullptr};
  size_t tb = 0, te = 0, ib = 0, ie = 0;
  unsigned char flags = 0;
  bool leaf() const { return (flags & 3) == 0; }
  int axis() const { return (flags & 3) - 1; }
};

struct Accel {
  std::vector<Mesh> meshes;
  std::vector<uint32_t> prim_mesh;
  std::vector<Xf> prim_xf;

  std::vector<F3> tri_v;  // 3 per triangle (tree order)
  std::vector<size_t> prim_idx, tri_idx;
  std::vector<BBoxF> tri_bb, inst_bb;

  std::vector<const Node *> inst_node;
  std::vector<Xf> inst_xf;
  std::vector<size_t> inst_prim;

  std::deque<Node> pool;
  Node *root = nullptr;
  size_t max_depth = 0;

  Node *alloc() {
    pool.emplace_back();
    return &pool.back();
  }

  void swap_tri(size_t a, size_t b) {
    for (int k = 0; k < 3; ++k) std::swap(tri_v[3 * a + k], tri_v[3 * b + k]);
    std::swap(prim_idx[a], prim_idx[b]);
    std::swap(tri_idx[a], tri_idx[b]);
    std::swap(tri_bb[a], tri_bb[b]);
  }
  void swap_inst(size_t a, size_t b) {
    std::swap(inst_node[a], inst_node[b]);
    std::swap(inst_xf[a], inst_xf[b]);
    std::swap(inst_prim[a], inst_prim[b]);
    std::swap(inst_bb[a], inst_bb[b]);
  }

  BBoxF make_bbox(size_t tb, size_t te, size_t ib, size_t ie) const {
    BBoxF b;
    for (size_t i = tb; i != te; ++i) b.unite(tri_bb[i]);
    for (size_t i = ib; i != ie; ++i) b.unite(inst_bb[i]);
    return b;
  }

  // TriangleAccelerator.cpp:505-617
  std::pair<int, double> best_split(const BBoxF &nb, size_t tb, size_t te, size_t ib, size_t ie,
                                    unsigned char mask) const {
    size_t nt = te - tb, ni = ie - ib;
    double best_pos[3] = {0.0, 0.0, 0.0}, best_cost[3] = {DBL_INF_, DBL_INF_, DBL_INF_};
    for (int axis = 0; axis < 3; ++axis) {
      if (mask & (1 << axis)) continue;
      BBoxD lb[MAX_SPLIT_TRIES], mb[MAX_SPLIT_TRIES], rb[MAX_SPLIT_TRIES];
      size_t ln[MAX_SPLIT_TRIES] = {0}, mn[MAX_SPLIT_TRIES] = {0}, rn[MAX_SPLIT_TRIES] = {0};
      size_t num_tries = smin(MAX_SPLIT_TRIES, 2 * nt + 2 * ni);
      // size_t / float is a float division in the reference
      double coef = (double)((float)num_tries / (nb.mx[axis] - nb.mn[axis]));
      for (size_t i = 0; i < nt + ni; ++i) {
        const BBoxF &bb = i < nt ? tri_bb[tb + i] : inst_bb[ib + i - nt];
        size_t lefts_begin = bin_of((double)(bb.mx[axis] - nb.mn[axis]) * coef + 1.0, num_tries);
        size_t rights_end = bin_of((double)(bb.mn[axis] - nb.mn[axis]) * coef + 1.0, num_tries);
        size_t cnt = 1;
        if (i >= nt) {
          const Node *sub = inst_node[ib + i - nt];
          cnt = sub->te - sub->tb;
        }
        if (rights_end > 0) {
          rb[rights_end - 1].unite(bb);
          rn[rights_end - 1] += cnt;
        }
        if (lefts_begin < num_tries) {
          lb[lefts_begin].unite(bb);
          ln[lefts_begin] += cnt;
        }
        for (size_t j = rights_end; j < lefts_begin; ++j) {
          mb[j].unite(bb);
          mn[j] += cnt;
        }
      }
      for (size_t i = 1; i < num_tries; ++i) {
        lb[i].unite(lb[i - 1]);
        ln[i] += ln[i - 1];
      }
      for (int i = (int)num_tries - 2; i >= 0; --i) {
        rb[i].unite(rb[i + 1]);
        rn[i] += rn[i + 1];
      }
      double inv_area = 1.0 / (double)nb.area();
      for (size_t i = 0; i < num_tries; ++i) {
        double p1 = smin(1.0, lb[i].area() * inv_area);
        double p2 = smin(1.0, mb[i].area() * inv_area);
        double p3 = smin(1.0, rb[i].area() * inv_area);
        double cost = (double)ln[i] * p1 + (double)mn[i] * p2 + (double)rn[i] * p3;
        if (ln[i] == 0 || rn[i] == 0) cost *= 0.8;
        if (cost < best_cost[axis]) {
          best_cost[axis] = cost;
          // float arithmetic in the reference: float + (size_t*float)/size_t
          float ext = nb.mx[axis] - nb.mn[axis];
          float v = (float)i * ext;
          v = v / (float)num_tries;
          best_pos[axis] = (double)(nb.mn[axis] + v);
        }
      }
    }
    if (best_cost[0] <= best_cost[1] && best_cost[0] <= best_cost[2]) return {0, best_pos[0]};
    if (best_cost[1] <= best_cost[0] && best_cost[1] <= best_cost[2]) return {1, best_pos[1]};
    return {2, best_pos[2]};
  }

  // TriangleAccelerator.cpp:621-729
  Node *build(size_t tb, size_t te, size_t ib, size_t ie, unsigned char mask, size_t depth) {
    Node *n = alloc();
    n->tb = tb;
    n->te = te;
    n->ib = ib;
    n->ie = ie;
    n->bbox = make_bbox(tb, te, ib, ie);
    if (depth > max_depth) max_depth = depth;
    if ((te - tb) + (ie - ib) <= MAX_TRIANGLES_IN_LEAF || mask == 7 || depth >= MAX_TREE_DEPTH) {
      n->flags = 0;
      return n;
    }
    std::pair<int, double> split = best_split(n->bbox, tb, te, ib, ie, mask);
    int axis = split.first;
    double pos = split.second;
    n->flags = (unsigned char)(axis + 1);

    size_t mid_b = tb, right_b = te, i = tb;
    while (i < right_b) {
      const BBoxF &bb = tri_bb[i];
      if (bb.mx[axis] < pos) {
        swap_tri(mid_b, i);
        ++mid_b;
        ++i;
      } else if (bb.mn[axis] > pos) {
        --right_b;
        swap_tri(i, right_b);
      } else {
        ++i;
      }
    }
    size_t imid_b = ib, iright_b = ie;
    i = ib;
    while (i < iright_b) {
      const BBoxF &bb = inst_bb[i];
      if (bb.mx[axis] < pos) {
        swap_inst(imid_b, i);
        ++imid_b;
        ++i;
      } else if (bb.mn[axis] > pos) {
        --iright_b;
        swap_inst(i, iright_b);
      } else {
        ++i;
      }
    }
    if (tb < mid_b || ib < imid_b) n->ch[0] = build(tb, mid_b, ib, imid_b, mask, depth + 1);
    if (mid_b < right_b || imid_b < iright_b)
      n->ch[1] = build(mid_b, right_b, imid_b, iright_b, (unsigned char)(mask | (1 << axis)), depth + 1);
    if (right_b < te || iright_b < ie) n->ch[2] = build(right_b, te, iright_b, ie, mask, depth + 1);
    return n;
  }

  void push_triangle(const F3 &a, const F3 &b, const F3 &c, size_t prim, size_t tri) {
    tri_v.push_back(a);
    tri_v.push_back(b);
    tri_v.push_back(c);
    prim_idx.push_back(prim);
    tri_idx.push_back(tri);
    BBoxF bb;
    bb.unite(a);
    bb.unite(b);
    bb.unite(c);
    tri_bb.push_back(bb);
  }

  // TriangleAccelerator.cpp:22-147
  void construct() {
    std::vector<std::vector<size_t>> users(meshes.size());
    for (size_t p = 0; p < prim_mesh.size(); ++p) users[prim_mesh[p]].push_back(p);

    for (size_t m = 0; m < meshes.size(); ++m) {
      if (users[m].size() <= 1) continue;
      const Mesh &mesh = meshes[m];
      size_t prev = tri_bb.size();
      for (size_t j = 0; j < mesh.tris.size() / 3; ++j)
        push_triangle(mesh.verts[mesh.tris[3 * j]], mesh.verts[mesh.tris[3 * j + 1]], mesh.verts[mesh.tris[3 * j + 2]],
                      NO_PRIM, j);
      const Node *sub = build(prev, tri_bb.size(), 0, 0, 0, 0);
      for (size_t k = 0; k < users[m].size(); ++k) {
        size_t p = users[m][k];
        const Xf &xf = prim_xf[p];
        BBoxF ib;
        for (int j = 0; j < 8; ++j) {
          F3 mn = {sub->bbox.mn[0], sub->bbox.mn[1], sub->bbox.mn[2]};
          F3 mx = {sub->bbox.mx[0], sub->bbox.mx[1], sub->bbox.mx[2]};
          F3 corner = {(j & 1) ? mn.x : mx.x, (j & 2) ? mn.y : mx.y, (j & 4) ? mn.z : mx.z};
          ib.unite(xf_point_f(xf.m, corner));
        }
        inst_prim.push_back(p);
        inst_node.push_back(sub);
        inst_xf.push_back(xf);
        inst_bb.push_back(ib);
      }
    }
    size_t prev = tri_bb.size();
    for (size_t m = 0; m < meshes.size(); ++m) {
      if (users[m].size() != 1) continue;
      const Mesh &mesh = meshes[m];
      size_t p = users[m][0];
      const Xf &xf = prim_xf[p];
      for (size_t j = 0; j < mesh.tris.size() / 3; ++j)
        push_triangle(xf_point_f(xf.m, mesh.verts[mesh.tris[3 * j]]), xf_point_f(xf.m, mesh.verts[mesh.tris[3 * j + 1]]),
                      xf_point_f(xf.m, mesh.verts[mesh.tris[3 * j + 2]]), p, j);
    }
    max_depth = 0;
    root = build(prev, tri_bb.size(), 0, inst_node.size(), 0, 0);
  }

  // Ray transform by an instance's world-to-mesh matrix (Transform.h:293-300).
  static Ray xf_ray(const double m[4][4], const Ray &r) {
    Ray t;
    t.o = xf_point(m, r.o);
    t.d = xf_vector(m, r.d);
    t.tmin = r.tmin;
    t.tmax = r.tmax;
    return t;
  }

  static bool slab_cull(const Node *n, const Ray &ray, const double inv[3]) {
    double tn1 = ((double)n->bbox.mn[0] - ray.o.x) * inv[0];
    double tf1 = ((double)n->bbox.mx[0] - ray.o.x) * inv[0];
    if (tn1 > tf1) std::swap(tn1, tf1);
    double tn2 = ((double)n->bbox.mn[1] - ray.o.y) * inv[1];
    double tf2 = ((double)n->bbox.mx[1] - ray.o.y) * inv[1];
    if (tn2 > tf2) std::swap(tn2, tf2);
    double tn3 = ((double)n->bbox.mn[2] - ray.o.z) * inv[2];
    double tf3 = ((double)n->bbox.mx[2] - ray.o.z) * inv[2];
    if (tn3 > tf3) std::swap(tn3, tf3);
    return ray.tmin > tf1 || tn1 > tf2 || tn2 > tf1 || tn3 > tf1 || ray.tmin > tf2 || tn1 > tf3 || tn2 > tf3 ||
           tn3 > tf2 || ray.tmin > tf3 || tn1 > ray.tmax || tn2 > ray.tmax || tn3 > ray.tmax;
  }

  // Möller–Trumbore in FP64 on float vertices; returns true and t if accepted in [tmin, tmax].
  bool tri_test(size_t i, const Ray &ray, double &t_out) const {
    V3 v0 = to_d(tri_v[3 * i]), v1 = to_d(tri_v[3 * i + 1]), v2 = to_d(tri_v[3 * i + 2]);
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 s1 = cross(ray.d, e2);
    double divisor = dot(s1, e1);
    double inv_div = 1.0 / divisor;
    V3 d = sub(ray.o, v0);
    double b1 = dot(d, s1) * inv_div;
    if (b1 > -DBL_EPS_ && b1 < (1.0 + DBL_EPS_)) {
      V3 s2 = cross(d, e1);
      double b2 = dot(ray.d, s2) * inv_div;
      if (b2 > -DBL_EPS_ && b1 + b2 < (1.0 + DBL_EPS_)) {
        double t = dot(e2, s2) * inv_div;
        if (t >= ray.tmin && t <= ray.tmax) {
          t_out = t;
          return true;
        }
      }
    }
    return false;
  }

  // TriangleAccelerator.cpp:154-293
  bool node_intersect(const Node *start, Ray &io_ray, size_t &o_prim, size_t &o_tri, uint64_t *pops,
                      uint64_t *tests) const {
    Ray ray = io_ray;
    double inv[3] = {1.0 / ray.d.x, 1.0 / ray.d.y, 1.0 / ray.d.z};
    const Node *todo[2 * MAX_TREE_DEPTH + 1];
    todo[0] = start;
    int sz = 1;
    bool hit = false, inst_hit = false;
    size_t tri_pos = 0;
    while (sz > 0) {
      const Node *n = todo[--sz];
      if (pops) ++*pops;
      if (slab_cull(n, ray, inv)) continue;
      if (!n->leaf()) {
        int axis = n->axis();
        if (ray.d[axis] > 0.0) {
          if (n->ch[2]) todo[sz++] = n->ch[2];
          if (n->ch[1]) todo[sz++] = n->ch[1];
          if (n->ch[0]) todo[sz++] = n->ch[0];
        } else {
          if (n->ch[0]) todo[sz++] = n->ch[0];
          if (n->ch[1]) todo[sz++] = n->ch[1];
          if (n->ch[2]) todo[sz++] = n->ch[2];
        }
      } else {
        for (size_t i = n->ib; i < n->ie; ++i) {
          Ray tr = xf_ray(inst_xf[i].inv, ray);
          size_t p2, t2;
          if (node_intersect(inst_node[i], tr, p2, t2, pops, tests)) {
            hit = true;
            inst_hit = true;
            ray.tmax = tr.tmax;
            o_tri = t2;
            o_prim = inst_prim[i];
          }
        }
        for (size_t i = n->tb; i < n->te; ++i) {
          if (tests) ++*tests;
          double t;
          if (tri_test(i, ray, t)) {
            ray.tmax = t;
            tri_pos = i;
            hit = true;
            inst_hit = false;
          }
        }
      }
    }
    io_ray.tmax = ray.tmax;
    if (inst_hit) return true;
    if (hit) {
      o_prim = prim_idx[tri_pos];
      o_tri = tri_idx[tri_pos];
      return true;
    }
    return false;
  }

  // TriangleAccelerator.cpp:354-462
  bool node_intersect_test(const Node *start, const Ray &ray) const {
    double inv[3] = {1.0 / ray.d.x, 1.0 / ray.d.y, 1.0 / ray.d.z};
    const Node *todo[2 * MAX_TREE_DEPTH + 1];
    todo[0] = start;
    int sz = 1;
    while (sz > 0) {
      const Node *n = todo[--sz];
      if (slab_cull(n, ray, inv)) continue;
      if (!n->leaf()) {
        int axis = n->axis();
        if (ray.d[axis] > 0.0) {
          if (n->ch[2]) todo[sz++] = n->ch[2];
          if (n->ch[1]) todo[sz++] = n->ch[1];
          if (n->ch[0]) todo[sz++] = n->ch[0];
        } else {
          if (n->ch[0]) todo[sz++] = n->ch[0];
          if (n->ch[1]) todo[sz++] = n->ch[1];
          if (n->ch[2]) todo[sz++] = n->ch[2];
        }
      } else {
        for (size_t i = n->ib; i < n->ie; ++i) {
          Ray tr = xf_ray(inst_xf[i].inv, ray);
          if (node_intersect_test(inst_node[i], tr)) return true;
        }
        for (size_t i = n->tb; i < n->te; ++i) {
          double t;
          if (tri_test(i, ray, t)) return true;
        }
      }
    }
    return false;
  }

  // TriangleMesh.cpp:414-437
  static bool intersection_point(const V3 v[3], V3 o, V3 d, double &b1, double &b2, double &t) {
    V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
    V3 s1 = cross(d, e2);
    double divisor = dot(s1, e1);
    if (divisor == 0.0) return false;
    double inv = 1.0 / divisor;
    V3 dd = sub(o, v[0]);
    b1 = dot(dd, s1) * inv;
    V3 s2 = cross(dd, e1);
    b2 = dot(d, s2) * inv;
    t = dot(e2, s2) * inv;
    return true;
  }

  // Intersect (TriangleAccelerator.cpp:295-352) with the parts of ComputeDifferentialGeometry
  // (TriangleMesh.cpp:439-500) that the hot path consumes.
  bool intersect(const Ray &in, orc_hit *h, orc_hit_geom *g, uint64_t *pops, uint64_t *tests) const {
    Ray ray = in;
    size_t prim = 0, tri = 0;
    h->prim = 0xffffffffu;
    h->tri = 0xffffffffu;
    h->t = DBL_INF_;
    h->b1 = 0.0;
    h->b2 = 0.0;
    if (!root || !node_intersect(root, ray, prim, tri, pops, tests)) return false;
    h->prim = (uint32_t)prim;
    h->tri = (uint32_t)tri;
    h->t = ray.tmax;
    const Mesh &mesh = meshes[prim_mesh[prim]];
    const Xf &xf = prim_xf[prim];
    Ray mr = xf_ray(xf.inv, in);
    V3 v[3] = {to_d(mesh.verts[mesh.tris[3 * tri]]), to_d(mesh.verts[mesh.tris[3 * tri + 1]]),
               to_d(mesh.verts[mesh.tris[3 * tri + 2]])};
    double b1 = 0, b2 = 0, tm = 0;
    intersection_point(v, mr.o, mr.d, b1, b2, tm);
    h->b1 = b1;
    h->b2 = b2;
    if (g) {
      V3 pm = add(mr.o, mul(mr.d, tm));
      V3 pw = xf_point(xf.m, pm);
      V3 gn = mesh.invert_normals ? cross(sub(v[2], v[0]), sub(v[1], v[0])) : cross(sub(v[1], v[0]), sub(v[2], v[0]));
      normalize(gn);
      V3 sn = gn;
      if (!mesh.normals.empty()) {
        double b0 = 1.0 - b1 - b2;
        V3 vn[3];
        for (int k = 0; k < 3; ++k) {
          vn[k] = to_d(mesh.normals[mesh.tris[3 * tri + k]]);
          if (mesh.invert_normals) vn[k] = mul(vn[k], -1.0);
        }
        sn = add(add(mul(vn[0], b0), mul(vn[1], b1)), mul(vn[2], b2));
        normalize(sn);
      }
      V3 gw = xf_normal(xf.inv, gn), sw = xf_normal(xf.inv, sn);
      V3 cr = cross(sub(v[1], v[0]), sub(v[2], v[0]));
      V3 pback = xf_point(xf.inv, pw);
      double dt = dot(cross(sub(v[0], pback), sub(v[1], v[0])), sub(v[2], v[0]));
      g->p[0] = pw.x; g->p[1] = pw.y; g->p[2] = pw.z;
      g->ng[0] = gw.x; g->ng[1] = gw.y; g->ng[2] = gw.z;
      g->ns[0] = sw.x; g->ns[1] = sw.y; g->ns[2] = sw.z;
      g->cross[0] = cr.x; g->cross[1] = cr.y; g->cross[2] = cr.z;
      g->dot = dt;
    }
    return true;
  }

  // CoreUtils.h:139-149
  double next_min_t(uint32_t prim, const orc_hit_geom &g, V3 dir) const {
    V3 dm = xf_vector(prim_xf[prim].inv, dir);
    double divisor = dot(dm, V3{g.cross[0], g.cross[1], g.cross[2]});
    if (divisor == 0.0) return 0.0;
    return smax(0.0, (g.dot * (1.0 / divisor)) + 1e-14);
  }

  // TriangleAccelerator.test.h:213-269 (first-wins on t ties, non-strict barycentric bounds)
  bool brute(const Ray &r, uint32_t &o_prim, uint32_t &o_tri, double &o_t) const {
    bool hit = false;
    o_t = std::numeric_limits<double>::max();
    for (size_t p = 0; p < prim_mesh.size(); ++p) {
      const Mesh &mesh = meshes[prim_mesh[p]];
      const Xf &xf = prim_xf[p];
      for (size_t j = 0; j < mesh.tris.size() / 3; ++j) {
        F3 a = mesh.verts[mesh.tris[3 * j]], b = mesh.verts[mesh.tris[3 * j + 1]], c = mesh.verts[mesh.tris[3 * j + 2]];
        if (!xf.identity) {
          a = xf_point_f(xf.m, a);
          b = xf_point_f(xf.m, b);
          c = xf_point_f(xf.m, c);
        }
        V3 v0 = to_d(a), v1 = to_d(b), v2 = to_d(c);
        V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
        V3 s1 = cross(r.d, e2);
        double divisor = dot(s1, e1);
        double inv = 1.0 / divisor;
        V3 d = sub(r.o, v0);
        double b1 = dot(d, s1) * inv;
        if (b1 < -DBL_EPS_ || b1 > (1.0 + DBL_EPS_) || divisor == 0.0) continue;
        V3 s2 = cross(d, e1);
        double b2 = dot(r.d, s2) * inv;
        if (b2 < -DBL_EPS_ || b1 + b2 > (1.0 + DBL_EPS_)) continue;
        double t = dot(e2, s2) * inv;
        if (t >= r.tmin && t <= r.tmax && t < o_t) {
          o_t = t;
          o_prim = (uint32_t)p;
          o_tri = (uint32_t)j;
          hit = true;
        }
      }
    }
    return hit;
  }
};

}  // namespace orc

using namespace orc;

static inline Ray load_ray(const double *r) { return Ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, r[6], r[7]}; }

extern "C" {

orc_accel *orc_accel_create(uint32_t n_meshes, const uint64_t *mesh_vert_off, const float *verts,
                            const float *normals, const uint64_t *mesh_tri_off, const uint32_t *tris,
                            const uint8_t *mesh_flags, uint32_t n_prims, const uint32_t *prim_mesh,
                            const double *prim_xf) {
  Accel *a = new Accel();
  a->meshes.resize(n_meshes);
  for (uint32_t m = 0; m < n_meshes; ++m) {
    Mesh &mesh = a->meshes[m];
    uint64_t vb = mesh_vert_off[m], ve = mesh_vert_off[m + 1];
    size_t nv = (size_t)(ve - vb);
    uint8_t fl = mesh_flags ? mesh_flags[m] : 0;
    mesh.verts.resize(nv);
    for (size_t i = 0; i < nv; ++i) mesh.verts[i] = {verts[3 * (vb + i)], verts[3 * (vb + i) + 1], verts[3 * (vb + i) + 2]};
    if ((fl & 2) && normals) {
      mesh.normals.resize(nv);
      for (size_t i = 0; i < nv; ++i) {
        // TriangleMesh stores Normalized() shading normals as floats (TriangleMesh.cpp:212-217)
        F3 nf{normals[3 * (vb + i)], normals[3 * (vb + i) + 1], normals[3 * (vb + i) + 2]};
        float l2 = nf.x * nf.x + nf.y * nf.y + nf.z * nf.z;
        if (l2 > 0.0f) {
          double inv = 1.0 / (double)std::sqrt(l2);
          nf = {(float)(nf.x * inv), (float)(nf.y * inv), (float)(nf.z * inv)};
        }
        mesh.normals[i] = nf;
      }
    }
    mesh.invert_normals = (fl & 1) != 0;
    // Stable removal of out-of-range and repeated-index triangles (TriangleMesh.cpp:134-147,180-186)
    for (uint64_t t = mesh_tri_off[m]; t < mesh_tri_off[m + 1]; ++t) {
      uint32_t i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
      bool bad = i0 >= nv || i1 >= nv || i2 >= nv || i0 == i1 || i0 == i2 || i1 == i2;
      if (bad) continue;
      mesh.tris.push_back(i0);
      mesh.tris.push_back(i1);
      mesh.tris.push_back(i2);
    }
  }
  a->prim_mesh.assign(prim_mesh, prim_mesh + n_prims);
  a->prim_xf.resize(n_prims);
  for (uint32_t p = 0; p < n_prims; ++p) {
    Xf &x = a->prim_xf[p];
    if (prim_xf) {
      std::memcpy(x.m, prim_xf + 32 * p, 16 * sizeof(double));
      std::memcpy(x.inv, prim_xf + 32 * p + 16, 16 * sizeof(double));
      x.identity = true;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          if (x.m[i][j] != (i == j ? 1.0 : 0.0) || x.inv[i][j] != (i == j ? 1.0 : 0.0)) x.identity = false;
    } else {
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) x.m[i][j] = x.inv[i][j] = (i == j) ? 1.0 : 0.0;
      x.identity = true;
    }
  }
  a->construct();
  return reinterpret_cast<orc_accel *>(a);
}

void orc_accel_destroy(orc_accel *h) { delete reinterpret_cast<Accel *>(h); }

void orc_accel_info(const orc_accel *h, uint64_t *out) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  out[0] = a->pool.size();
  out[1] = a->tri_bb.size();
  out[2] = a->max_depth;
  out[3] = a->inst_node.size();
  size_t leaves = 0;
  for (const Node &n : a->pool) leaves += n.leaf() ? 1 : 0;
  out[4] = leaves;
}

void orc_accel_bounds(const orc_accel *h, double *out6) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  for (int i = 0; i < 3; ++i) {
    out6[i] = a->root ? a->root->bbox.mn[i] : 0.0;
    out6[3 + i] = a->root ? a->root->bbox.mx[i] : 0.0;
  }
}

// Depth-first pre-order dump of the top-level tree (children in left, middle, right order).
// Per node: bbox (6 floats); meta = [flags, child mask, tri_begin, tri_end, inst_begin, inst_end].
// Triangle order: (prim, tri) per tree position.
uint64_t orc_accel_dump(const orc_accel *h, float *bbox, uint64_t *meta, uint32_t *tri_prim, uint32_t *tri_tri) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  uint64_t k = 0;
  if (a->root) {
    std::vector<const Node *> st{a->root};
    while (!st.empty()) {
      const Node *n = st.back();
      st.pop_back();
      if (bbox)
        for (int i = 0; i < 3; ++i) {
          bbox[6 * k + i] = n->bbox.mn[i];
          bbox[6 * k + 3 + i] = n->bbox.mx[i];
        }
      if (meta) {
        uint64_t mask = (n->ch[0] ? 1 : 0) | (n->ch[1] ? 2 : 0) | (n->ch[2] ? 4 : 0);
        meta[6 * k + 0] = n->flags & 3;
        meta[6 * k + 1] = n->leaf() ? 0 : mask;
        // ranges are reported for leaves only (internal nodes cover the union of their children)
        meta[6 * k + 2] = n->leaf() ? n->tb : 0;
        meta[6 * k + 3] = n->leaf() ? n->te : 0;
        meta[6 * k + 4] = n->leaf() ? n->ib : 0;
        meta[6 * k + 5] = n->leaf() ? n->ie : 0;
      }
      ++k;
      if (!n->leaf())
        for (int c = 2; c >= 0; --c)
          if (n->ch[c]) st.push_back(n->ch[c]);
    }
  }
  if (tri_prim)
    for (size_t i = 0; i < a->tri_bb.size(); ++i) {
      tri_prim[i] = a->prim_idx[i] == NO_PRIM ? 0xffffffffu : (uint32_t)a->prim_idx[i];
      tri_tri[i] = (uint32_t)a->tri_idx[i];
    }
  return k;
}

void orc_closest_hit(const orc_accel *h, uint64_t n, const double *rays, orc_hit *hits, orc_hit_geom *geom,
                     uint32_t *pops, uint32_t *tests, int nthreads) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t p = 0, t = 0;
    a->intersect(load_ray(rays + 8 * i), &hits[i], geom ? &geom[i] : nullptr, pops ? &p : nullptr,
                 tests ? &t : nullptr);
    if (pops) pops[i] = (uint32_t)p;
    if (tests) tests[i] = (uint32_t)t;
  }
}

void orc_any_hit(const orc_accel *h, uint64_t n, const double *rays, uint8_t *out, int nthreads) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    Ray r = load_ray(rays + 8 * i);
    out[i] = (a->root && a->node_intersect_test(a->root, r)) ? 1 : 0;
  }
}

void orc_brute_closest(const orc_accel *h, uint64_t n, const double *rays, orc_hit *hits, int nthreads) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint32_t p = 0xffffffffu, t = 0xffffffffu;
    double tt;
    bool hit = a->brute(load_ray(rays + 8 * i), p, t, tt);
    hits[i].prim = hit ? p : 0xffffffffu;
    hits[i].tri = hit ? t : 0xffffffffu;
    hits[i].t = hit ? tt : DBL_INF_;
    hits[i].b1 = 0.0;
    hits[i].b2 = 0.0;
  }
}

double orc_next_min_t(const orc_accel *h, uint32_t prim, const orc_hit_geom *g, const double *dir3) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  return a->next_min_t(prim, *g, V3{dir3[0], dir3[1], dir3[2]});
}

int orc_accel_intersect_one(const orc_accel *h, const double *ray, orc_hit *hit, orc_hit_geom *geom) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  return a->intersect(load_ray(ray), hit, geom, nullptr, nullptr) ? 1 : 0;
}

uint32_t orc_accel_prim_count(const orc_accel *h) {
  return (uint32_t) reinterpret_cast<const Accel *>(h)->prim_mesh.size();
}

uint32_t orc_accel_mesh_tri_count(const orc_accel *h, uint32_t prim) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  return (uint32_t)(a->meshes[a->prim_mesh[prim]].tris.size() / 3);
}

int orc_accel_mesh_inverted(const orc_accel *h, uint32_t prim) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  return a->meshes[a->prim_mesh[prim]].invert_normals ? 1 : 0;
}

void orc_accel_mesh_tri(const orc_accel *h, uint32_t prim, uint32_t tri, float *v9) {
  const Accel *a = reinterpret_cast<const Accel *>(h);
  const Mesh &m = a->meshes[a->prim_mesh[prim]];
  for (int k = 0; k < 3; ++k) {
    F3 p = m.verts[m.tris[3 * tri + k]];
    v9[3 * k] = p.x;
    v9[3 * k + 1] = p.y;
    v9[3 * k + 2] = p.z;
  }
}

}  // extern "C"
