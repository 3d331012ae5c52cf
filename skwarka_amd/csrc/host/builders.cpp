// Host builders for the device layouts.
//
// Triangle tree: the construction of Skwarka's TriangleAccelerator
// (Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:22-147, 470-729), rebuilt for a flat
// device layout:
//  * the partition moves 28-byte (bbox, triangle reference) records instead of whole triangles;
//  * subtrees are built on concurrent host threads (their ranges are disjoint, so the result is
//    the same as a serial build) and then laid out children-block by children-block;
//  * mesh order is the caller's mesh order (the reference walks a std::map keyed by mesh address);
//  * a NaN bin position (zero node extent on an axis) maps to num_tries, the value the x86-64 SSE2
//    build of the reference produces for that undefined conversion.
// Split arithmetic keeps the reference's types exactly: float bin coefficient (size_t / float),
// double cost terms, float split position (float + size_t*float/size_t).
//
// Photon kd-tree: KDTree<Photon>::_Build (Source/RayLibs/Raytracer/Core/KDTree.h:427-473):
// longest-axis split of the double bbox, std::nth_element median, median swapped to the front.
#include "builders.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <memory>
#include <thread>

namespace skw {
namespace {

constexpr size_t kMaxLeaf = 4;      // TriangleAccelerator.h:128
constexpr size_t kMaxTries = 200;   // TriangleAccelerator.h:131
constexpr size_t kMaxDepth = 200;   // TriangleAccelerator.h:134
constexpr uint32_t kNoPrim = 0xFFFFFFFFu;
constexpr size_t kParallelMin = 16384;  // spawn a host thread for subtrees larger than this

template <class T> inline T fmin_std(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T fmax_std(T a, T b) { return (a < b) ? b : a; }

struct BoxF {
  float mn[3], mx[3];
  void clear() {
    for (int i = 0; i < 3; ++i) {
      mn[i] = std::numeric_limits<float>::infinity();
      mx[i] = -std::numeric_limits<float>::infinity();
    }
  }
  void add(const BoxF &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = fmin_std(mn[i], b.mn[i]);
      mx[i] = fmax_std(mx[i], b.mx[i]);
    }
  }
  void add_point(const float *p) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = fmin_std(mn[i], p[i]);
      mx[i] = fmax_std(mx[i], p[i]);
    }
  }
  float area() const {
    float x = std::fabs(mx[0] - mn[0]), y = std::fabs(mx[1] - mn[1]), z = std::fabs(mx[2] - mn[2]);
    return 2.0f * (x * y + x * z + y * z);
  }
};

struct BoxD {
  double mn[3], mx[3];
  void clear() {
    for (int i = 0; i < 3; ++i) {
      mn[i] = std::numeric_limits<double>::infinity();
      mx[i] = -std::numeric_limits<double>::infinity();
    }
  }
  void add(const BoxF &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = fmin_std(mn[i], (double)b.mn[i]);
      mx[i] = fmax_std(mx[i], (double)b.mx[i]);
    }
  }
  void add(const BoxD &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = fmin_std(mn[i], b.mn[i]);
      mx[i] = fmax_std(mx[i], b.mx[i]);
    }
  }
  double area() const {
    double x = std::fabs(mx[0] - mn[0]), y = std::fabs(mx[1] - mn[1]), z = std::fabs(mx[2] - mn[2]);
    return 2.0 * (x * y + x * z + y * z);
  }
};

inline size_t bin_index(double v, size_t num_tries) {
  if (!(v < 18446744073709551616.0)) return num_tries;
  return fmin_std(num_tries, (size_t)v);
}

struct BNode {
  BoxF box;
  std::unique_ptr<BNode> ch[3];
  uint64_t tb = 0, te = 0, ib = 0, ie = 0;
  uint8_t flags = 0;  // 0 leaf, axis + 1
  uint32_t depth = 0;
};

// Transform of a float point by a double matrix, as Transform::operator()(Point3D<float>)
// (Source/RayLibs/Math/Transform.h:223-238).
void xform_point_f(const double m[16], const float *p, float *out) {
  double x = p[0], y = p[1], z = p[2];
  float xp = (float)(m[0] * x + m[1] * y + m[2] * z + m[3]);
  float yp = (float)(m[4] * x + m[5] * y + m[6] * z + m[7]);
  float zp = (float)(m[8] * x + m[9] * y + m[10] * z + m[11]);
  float wp = (float)(m[12] * x + m[13] * y + m[14] * z + m[15]);
  if (wp == 1.0f) {
    out[0] = xp;
    out[1] = yp;
    out[2] = zp;
    return;
  }
  double inv = 1.0 / (double)wp;
  out[0] = (float)(xp * inv);
  out[1] = (float)(yp * inv);
  out[2] = (float)(zp * inv);
}

class TriTreeBuilder {
 public:
  // Unique triangles (world space for plain meshes, mesh space for instanced meshes).
  std::vector<float> tv;        // 9 per unique triangle
  std::vector<uint32_t> tprim;  // per unique triangle
  std::vector<uint32_t> ttri;

  // Partition arrays (positions in the final triangle order).
  std::vector<BoxF> pbox;
  std::vector<uint32_t> pref;

  std::vector<BoxF> ibox;
  std::vector<uint32_t> iref;  // index into inst_* arrays
  std::vector<uint64_t> inst_count;
  std::vector<const BNode *> inst_root;
  std::vector<uint32_t> inst_prim;

  std::atomic<int> threads{0};
  int max_threads = 1;

  BoxF node_box(uint64_t tb, uint64_t te, uint64_t ib, uint64_t ie) const {
    BoxF b;
    b.clear();
    for (uint64_t i = tb; i != te; ++i) b.add(pbox[i]);
    for (uint64_t i = ib; i != ie; ++i) b.add(ibox[i]);
    return b;
  }

  void axis_cost(const BoxF &nb, uint64_t tb, uint64_t te, uint64_t ib, uint64_t ie, int axis, double &best_cost,
                 double &best_pos) const {
    uint64_t nt = te - tb, ni = ie - ib;
    std::vector<BoxD> lb(kMaxTries), mb(kMaxTries), rb(kMaxTries);
    for (size_t i = 0; i < kMaxTries; ++i) {
      lb[i].clear();
      mb[i].clear();
      rb[i].clear();
    }
    size_t ln[kMaxTries] = {0}, mn[kMaxTries] = {0}, rn[kMaxTries] = {0};
    size_t num_tries = fmin_std(kMaxTries, (size_t)(2 * nt + 2 * ni));
    double coef = (double)((float)num_tries / (nb.mx[axis] - nb.mn[axis]));
    float lo = nb.mn[axis];
    for (uint64_t i = 0; i < nt + ni; ++i) {
      const BoxF &bb = i < nt ? pbox[tb + i] : ibox[ib + i - nt];
      size_t lefts_begin = bin_index((double)(bb.mx[axis] - lo) * coef + 1.0, num_tries);
      size_t rights_end = bin_index((double)(bb.mn[axis] - lo) * coef + 1.0, num_tries);
      size_t cnt = i < nt ? 1 : (size_t)inst_count[iref[ib + i - nt]];
      if (rights_end > 0) {
        rb[rights_end - 1].add(bb);
        rn[rights_end - 1] += cnt;
      }
      if (lefts_begin < num_tries) {
        lb[lefts_begin].add(bb);
        ln[lefts_begin] += cnt;
      }
      for (size_t j = rights_end; j < lefts_begin; ++j) {
        mb[j].add(bb);
        mn[j] += cnt;
      }
    }
    for (size_t i = 1; i < num_tries; ++i) {
      lb[i].add(lb[i - 1]);
      ln[i] += ln[i - 1];
    }
    for (int i = (int)num_tries - 2; i >= 0; --i) {
      rb[i].add(rb[i + 1]);
      rn[i] += rn[i + 1];
    }
    double inv_area = 1.0 / (double)nb.area();
    for (size_t i = 0; i < num_tries; ++i) {
      double p1 = fmin_std(1.0, lb[i].area() * inv_area);
      double p2 = fmin_std(1.0, mb[i].area() * inv_area);
      double p3 = fmin_std(1.0, rb[i].area() * inv_area);
      double cost = (double)ln[i] * p1 + (double)mn[i] * p2 + (double)rn[i] * p3;
      if (ln[i] == 0 || rn[i] == 0) cost *= 0.8;
      if (cost < best_cost) {
        best_cost = cost;
        float ext = nb.mx[axis] - nb.mn[axis];
        float step = ((float)i * ext) / (float)num_tries;
        best_pos = (double)(nb.mn[axis] + step);
      }
    }
  }

  std::pair<int, double> best_split(const BoxF &nb, uint64_t tb, uint64_t te, uint64_t ib, uint64_t ie,
                                    uint8_t mask) {
    double cost[3] = {std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                      std::numeric_limits<double>::infinity()};
    double pos[3] = {0.0, 0.0, 0.0};
    bool par = (te - tb) + (ie - ib) > 4 * kParallelMin;
    std::thread th[3];
    for (int a = 0; a < 3; ++a) {
      if (mask & (1 << a)) continue;
      if (par)
        th[a] = std::thread([&, a] { axis_cost(nb, tb, te, ib, ie, a, cost[a], pos[a]); });
      else
        axis_cost(nb, tb, te, ib, ie, a, cost[a], pos[a]);
    }
    for (int a = 0; a < 3; ++a)
      if (th[a].joinable()) th[a].join();
    if (cost[0] <= cost[1] && cost[0] <= cost[2]) return {0, pos[0]};
    if (cost[1] <= cost[0] && cost[1] <= cost[2]) return {1, pos[1]};
    return {2, pos[2]};
  }

  std::unique_ptr<BNode> build(uint64_t tb, uint64_t te, uint64_t ib, uint64_t ie, uint8_t mask, uint32_t depth) {
    std::unique_ptr<BNode> n(new BNode());
    n->tb = tb;
    n->te = te;
    n->ib = ib;
    n->ie = ie;
    n->depth = depth;
    n->box = node_box(tb, te, ib, ie);
    if ((te - tb) + (ie - ib) <= kMaxLeaf || mask == 7 || depth >= kMaxDepth) {
      n->flags = 0;
      return n;
    }
    std::pair<int, double> sp = best_split(n->box, tb, te, ib, ie, mask);
    int axis = sp.first;
    double split = sp.second;
    n->flags = (uint8_t)(axis + 1);

    uint64_t mid = tb, right = te, i = tb;
    while (i < right) {
      const BoxF &bb = pbox[i];
      if (bb.mx[axis] < split) {
        std::swap(pbox[mid], pbox[i]);
        std::swap(pref[mid], pref[i]);
        ++mid;
        ++i;
      } else if (bb.mn[axis] > split) {
        --right;
        std::swap(pbox[i], pbox[right]);
        std::swap(pref[i], pref[right]);
      } else {
        ++i;
      }
    }
    uint64_t imid = ib, iright = ie;
    i = ib;
    while (i < iright) {
      const BoxF &bb = ibox[i];
      if (bb.mx[axis] < split) {
        std::swap(ibox[imid], ibox[i]);
        std::swap(iref[imid], iref[i]);
        ++imid;
        ++i;
      } else if (bb.mn[axis] > split) {
        --iright;
        std::swap(ibox[i], ibox[iright]);
        std::swap(iref[i], iref[iright]);
      } else {
        ++i;
      }
    }
    struct Range {
      bool present;
      uint64_t tb, te, ib, ie;
      uint8_t mask;
    } r[3] = {{tb < mid || ib < imid, tb, mid, ib, imid, mask},
              {mid < right || imid < iright, mid, right, imid, iright, (uint8_t)(mask | (1 << axis))},
              {right < te || iright < ie, right, te, iright, ie, mask}};
    std::future<std::unique_ptr<BNode>> fut[3];
    for (int c = 0; c < 3; ++c) {
      if (!r[c].present) continue;
      uint64_t sz = (r[c].te - r[c].tb) + (r[c].ie - r[c].ib);
      if (c < 2 && sz > kParallelMin && threads.fetch_add(1) < max_threads) {
        Range rc = r[c];
        fut[c] = std::async(std::launch::async, [this, rc, depth] {
          auto res = build(rc.tb, rc.te, rc.ib, rc.ie, rc.mask, depth + 1);
          threads.fetch_sub(1);
          return res;
        });
      } else {
        if (c < 2 && sz > kParallelMin) threads.fetch_sub(1);
        n->ch[c] = build(r[c].tb, r[c].te, r[c].ib, r[c].ie, r[c].mask, depth + 1);
      }
    }
    for (int c = 0; c < 3; ++c)
      if (fut[c].valid()) n->ch[c] = fut[c].get();
    return n;
  }
};

uint32_t subtree_depth(const BNode *n) {
  uint32_t d = n->depth;
  for (int c = 0; c < 3; ++c)
    if (n->ch[c]) d = std::max(d, subtree_depth(n->ch[c].get()));
  return d;
}

// Lays out a subtree: the node at index idx, its children as one block appended at the end.
void place(const BNode *n, uint32_t idx, TreeLayout &out, bool with_inst) {
  DevNode &d = out.nodes[idx];
  for (int i = 0; i < 3; ++i) {
    d.bmin[i] = n->box.mn[i];
    d.bmax[i] = n->box.mx[i];
  }
  if (with_inst) {
    out.inst_range[2 * idx] = (uint32_t)n->ib;
    out.inst_range[2 * idx + 1] = (uint32_t)n->ie;
  }
  if (n->flags == 0) {
    d.a = (uint32_t)n->tb;
    uint32_t cnt = (uint32_t)(n->te - n->tb);
    d.b = (cnt << 3) | ((n->ie > n->ib) ? 4u : 0u);
    out.leaves++;
    return;
  }
  uint32_t mask = 0, nch = 0;
  for (int c = 0; c < 3; ++c)
    if (n->ch[c]) {
      mask |= 1u << c;
      ++nch;
    }
  uint32_t first = (uint32_t)out.nodes.size();
  d.a = first;
  d.b = (uint32_t)n->flags | (mask << 2);
  out.nodes.resize(out.nodes.size() + nch);
  if (with_inst) out.inst_range.resize(2 * out.nodes.size());
  uint32_t k = 0;
  for (int c = 0; c < 3; ++c)
    if (n->ch[c]) place(n->ch[c].get(), first + k++, out, with_inst);
}

}  // namespace

// Checks the description, cleans the meshes (TriangleMesh.cpp:174-229: out-of-range and repeated
// vertex indices dropped, the rest renumbered in order) and reads the primitives.
static bool read_scene(const skw_scene_desc &desc, TreeLayout &out, std::string &err) {
  if (desc.n_meshes > 0 && (!desc.mesh_vert_offset || !desc.mesh_tri_offset)) {
    err = "mesh offsets are NULL";
    return false;
  }
  if (desc.n_prims > 0 && !desc.prim_mesh) {
    err = "prim_mesh is NULL";
    return false;
  }
  out = TreeLayout();
  out.meshes.resize(desc.n_meshes);
  for (uint32_t m = 0; m < desc.n_meshes; ++m) {
    HostMesh &mesh = out.meshes[m];
    uint64_t vb = desc.mesh_vert_offset[m], ve = desc.mesh_vert_offset[m + 1];
    uint64_t tb = desc.mesh_tri_offset[m], te = desc.mesh_tri_offset[m + 1];
    if (ve < vb || te < tb) {
      err = "mesh offsets decrease";
      return false;
    }
    if (ve > vb && !desc.verts) {
      err = "verts is NULL";
      return false;
    }
    if (te > tb && !desc.tris) {
      err = "tris is NULL";
      return false;
    }
    uint64_t nv = ve - vb;
    mesh.verts.assign(desc.verts + 3 * vb, desc.verts + 3 * ve);
    const uint8_t mflags = desc.mesh_flags ? desc.mesh_flags[m] : 0;
    if ((mflags & SKW_MESH_SHADING_NORMALS) && !desc.normals) {
      err = "a mesh asks for shading normals but normals is NULL";
      return false;
    }
    if (mflags & SKW_MESH_SHADING_NORMALS) {
      mesh.normals.resize(3 * nv);
      for (uint64_t i = 0; i < nv; ++i) {
        // Vector3D<float>::Normalized(): float length, double reciprocal, float result.
        const float *n = desc.normals + 3 * (vb + i);
        float l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        float *o = &mesh.normals[3 * i];
        o[0] = n[0];
        o[1] = n[1];
        o[2] = n[2];
        if (l2 > 0.0f) {
          double inv = 1.0 / (double)std::sqrt(l2);
          for (int c = 0; c < 3; ++c) o[c] = (float)(o[c] * inv);
        }
      }
    }
    mesh.invert_normals = (mflags & SKW_MESH_INVERT_NORMALS) != 0;
    mesh.tris.reserve(3 * (te - tb));
    for (uint64_t t = tb; t < te; ++t) {
      uint32_t i0 = desc.tris[3 * t], i1 = desc.tris[3 * t + 1], i2 = desc.tris[3 * t + 2];
      if (i0 >= nv || i1 >= nv || i2 >= nv || i0 == i1 || i0 == i2 || i1 == i2) continue;
      mesh.tris.push_back(i0);
      mesh.tris.push_back(i1);
      mesh.tris.push_back(i2);
    }
  }
  out.prims.resize(desc.n_prims);
  for (uint32_t p = 0; p < desc.n_prims; ++p) {
    HostPrim &hp = out.prims[p];
    if (desc.prim_mesh[p] >= desc.n_meshes) {
      err = "prim_mesh index out of range";
      return false;
    }
    hp.mesh = desc.prim_mesh[p];
    hp.identity = true;
    for (int i = 0; i < 16; ++i) {
      double id = (i % 5 == 0) ? 1.0 : 0.0;
      hp.m2w[i] = desc.prim_xform ? desc.prim_xform[32 * p + i] : id;
      hp.w2m[i] = desc.prim_xform ? desc.prim_xform[32 * p + 16 + i] : id;
      if (hp.m2w[i] != id || hp.w2m[i] != id) hp.identity = false;
    }
  }

  return true;
}

bool build_triangle_tree(const skw_scene_desc &desc, TreeLayout &out, std::string &err) {
  if (!read_scene(desc, out, err)) return false;
  TriTreeBuilder b;
  b.max_threads = std::max(1, (int)std::thread::hardware_concurrency());
  std::vector<std::vector<uint32_t>> users(desc.n_meshes);
  for (uint32_t p = 0; p < desc.n_prims; ++p) users[desc.prim_mesh[p]].push_back(p);

  auto push_tri = [&](const float *a, const float *bv, const float *c, uint32_t prim, uint32_t tri) {
    uint32_t id = (uint32_t)b.tprim.size();
    b.tv.insert(b.tv.end(), a, a + 3);
    b.tv.insert(b.tv.end(), bv, bv + 3);
    b.tv.insert(b.tv.end(), c, c + 3);
    b.tprim.push_back(prim);
    b.ttri.push_back(tri);
    BoxF bb;
    bb.clear();
    bb.add_point(a);
    bb.add_point(bv);
    bb.add_point(c);
    b.pbox.push_back(bb);
    b.pref.push_back(id);
  };

  std::vector<std::unique_ptr<BNode>> inst_trees;
  std::vector<uint32_t> inst_tree_of;  // per instance: index into inst_trees
  std::vector<const double *> inst_w2m;
  for (uint32_t m = 0; m < desc.n_meshes; ++m) {
    if (users[m].size() <= 1) continue;
    const HostMesh &mesh = out.meshes[m];
    uint64_t prev = b.pbox.size();
    for (size_t j = 0; j < mesh.tris.size() / 3; ++j)
      push_tri(&mesh.verts[3 * mesh.tris[3 * j]], &mesh.verts[3 * mesh.tris[3 * j + 1]],
               &mesh.verts[3 * mesh.tris[3 * j + 2]], kNoPrim, (uint32_t)j);
    inst_trees.push_back(b.build(prev, b.pbox.size(), 0, 0, 0, 0));
    const BNode *sub = inst_trees.back().get();
    for (uint32_t p : users[m]) {
      const HostPrim &hp = out.prims[p];
      BoxF ib;
      ib.clear();
      for (int j = 0; j < 8; ++j) {
        float corner[3] = {(j & 1) ? sub->box.mn[0] : sub->box.mx[0], (j & 2) ? sub->box.mn[1] : sub->box.mx[1],
                           (j & 4) ? sub->box.mn[2] : sub->box.mx[2]};
        float w[3];
        xform_point_f(hp.m2w, corner, w);
        ib.add_point(w);
      }
      uint32_t inst = (uint32_t)b.inst_prim.size();
      b.inst_prim.push_back(p);
      b.inst_root.push_back(sub);
      b.inst_count.push_back(sub->te - sub->tb);
      b.ibox.push_back(ib);
      b.iref.push_back(inst);
      inst_tree_of.push_back((uint32_t)inst_trees.size() - 1);
    }
  }
  uint64_t prev = b.pbox.size();
  for (uint32_t m = 0; m < desc.n_meshes; ++m) {
    if (users[m].size() != 1) continue;
    const HostMesh &mesh = out.meshes[m];
    uint32_t p = users[m][0];
    const HostPrim &hp = out.prims[p];
    for (size_t j = 0; j < mesh.tris.size() / 3; ++j) {
      float w[3][3];
      for (int k = 0; k < 3; ++k) xform_point_f(hp.m2w, &mesh.verts[3 * mesh.tris[3 * j + k]], w[k]);
      push_tri(w[0], w[1], w[2], p, (uint32_t)j);
    }
  }
  if (b.pbox.size() >= (1ull << 29) || b.inst_prim.size() >= (1ull << 31)) {
    err = "too many triangles for the 29-bit leaf count";
    return false;
  }
  std::unique_ptr<BNode> root = b.build(prev, b.pbox.size(), 0, b.inst_prim.size(), 0, 0);

  // Lay out: top-level tree first (root at 0), then each instanced subtree.
  bool with_inst = !b.inst_prim.empty();
  out.nodes.resize(1);
  if (with_inst) out.inst_range.assign(2, 0);
  place(root.get(), 0, out, with_inst);
  out.root = 0;
  out.max_depth = subtree_depth(root.get());
  uint32_t inst_depth = 0;
  std::vector<uint32_t> inst_tree_root(inst_trees.size());
  for (size_t t = 0; t < inst_trees.size(); ++t) {
    uint32_t idx = (uint32_t)out.nodes.size();
    out.nodes.resize(idx + 1);
    if (with_inst) out.inst_range.resize(2 * out.nodes.size());
    place(inst_trees[t].get(), idx, out, with_inst);
    inst_tree_root[t] = idx;
    inst_depth = std::max(inst_depth, subtree_depth(inst_trees[t].get()));
  }
  out.stack_need = out.max_depth + 1 + (with_inst ? inst_depth + 1 : 0);
  for (int i = 0; i < 3; ++i) {
    out.bounds[i] = root->box.mn[i];
    out.bounds[3 + i] = root->box.mx[i];
  }
  if (b.pbox.empty() && b.inst_prim.empty()) {
    // An empty scene: the root box stays (+inf, -inf), so every ray misses.
    out.leaves = 1;
  }

  out.tris.resize(b.pref.size());
  for (size_t pos = 0; pos < b.pref.size(); ++pos) {
    uint32_t id = b.pref[pos];
    DevTri &t = out.tris[pos];
    std::memcpy(t.v0, &b.tv[9 * (size_t)id], 3 * sizeof(float));
    std::memcpy(t.v1, &b.tv[9 * (size_t)id + 3], 3 * sizeof(float));
    std::memcpy(t.v2, &b.tv[9 * (size_t)id + 6], 3 * sizeof(float));
    t.prim = b.tprim[id];
    t.tri = b.ttri[id];
    t.pad = 0;
  }
  out.insts.resize(b.inst_prim.size());
  for (size_t pos = 0; pos < b.iref.size(); ++pos) {
    uint32_t inst = b.iref[pos];
    DevInst &di = out.insts[pos];
    std::memcpy(di.w2m, out.prims[b.inst_prim[inst]].w2m, 16 * sizeof(double));
    di.root = inst_tree_root[inst_tree_of[inst]];
    di.prim = b.inst_prim[inst];
    di.pad[0] = di.pad[1] = 0;
  }
  return true;
}

bool read_plain_scene(const skw_scene_desc &desc, TreeLayout &out, std::vector<uint64_t> &goff,
                      std::vector<uint32_t> &user, bool &instanced, std::string &err) {
  instanced = false;
  goff.clear();
  user.clear();
  if (!read_scene(desc, out, err)) return false;
  std::vector<uint32_t> users(desc.n_meshes, 0);
  user.assign(desc.n_meshes, 0);
  for (uint32_t p = 0; p < desc.n_prims; ++p) {
    users[desc.prim_mesh[p]]++;
    user[desc.prim_mesh[p]] = p;
  }
  for (uint32_t m = 0; m < desc.n_meshes; ++m)
    if (users[m] > 1) {
      instanced = true;
      user.clear();
      return true;
    }
  // the builder's order: meshes with one primitive, in mesh order, their triangles in order
  goff.assign(desc.n_meshes + 1, 0);
  for (uint32_t m = 0; m < desc.n_meshes; ++m)
    goff[m + 1] = goff[m] + (users[m] == 1 ? out.meshes[m].tris.size() / 3 : 0);
  if (goff.back() >= (1ull << 29)) {
    err = "too many triangles for the 29-bit leaf count";
    return false;
  }
  return true;
}

uint64_t export_tree(const TreeLayout &t, float *bbox, uint64_t *meta, uint32_t *tri_prim, uint32_t *tri_tri) {
  uint64_t k = 0;
  bool with_inst = !t.inst_range.empty();
  std::vector<uint32_t> st{t.root};
  while (!st.empty()) {
    uint32_t idx = st.back();
    st.pop_back();
    const DevNode &n = t.nodes[idx];
    if (bbox)
      for (int i = 0; i < 3; ++i) {
        bbox[6 * k + i] = n.bmin[i];
        bbox[6 * k + 3 + i] = n.bmax[i];
      }
    bool leaf = (n.b & 3) == 0;
    uint32_t mask = leaf ? 0 : (n.b >> 2) & 7;
    if (meta) {
      meta[6 * k + 0] = n.b & 3;
      meta[6 * k + 1] = mask;
      meta[6 * k + 2] = leaf ? n.a : 0;
      meta[6 * k + 3] = leaf ? n.a + (n.b >> 3) : 0;
      meta[6 * k + 4] = (leaf && with_inst) ? t.inst_range[2 * idx] : 0;
      meta[6 * k + 5] = (leaf && with_inst) ? t.inst_range[2 * idx + 1] : 0;
    }
    ++k;
    if (!leaf) {
      uint32_t nch = __builtin_popcount(mask);
      for (int c = (int)nch - 1; c >= 0; --c) st.push_back(n.a + (uint32_t)c);
    }
  }
  if (tri_prim)
    for (size_t i = 0; i < t.tris.size(); ++i) {
      tri_prim[i] = t.tris[i].prim;
      tri_tri[i] = t.tris[i].tri;
    }
  return k;
}

namespace {

struct KdItem {
  float p[3];
  uint32_t id;
};

struct KdBuilder {
  std::vector<KdItem> items;
  std::vector<uint32_t> word;
  std::vector<float> split;
  std::atomic<int> threads{0};
  int max_threads = 1;
  std::atomic<uint32_t> max_depth{0};

  void build(size_t b, size_t e, uint32_t depth) {
    uint32_t md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    if (b + 1 == e) {
      word[b] = (word[b] & ~3u) + 3u;
      return;
    }
    double mn[3], mx[3];
    for (int i = 0; i < 3; ++i) {
      mn[i] = std::numeric_limits<double>::infinity();
      mx[i] = -std::numeric_limits<double>::infinity();
    }
    for (size_t i = b; i < e; ++i)
      for (int c = 0; c < 3; ++c) {
        double v = items[i].p[c];
        mn[c] = fmin_std(mn[c], v);
        mx[c] = fmax_std(mx[c], v);
      }
    int axis = 0;
    if (mx[1] - mn[1] > mx[axis] - mn[axis]) axis = 1;
    if (mx[2] - mn[2] > mx[axis] - mn[axis]) axis = 2;
    size_t mid = (b + e) / 2;
    std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e,
                     [axis](const KdItem &x, const KdItem &y) { return x.p[axis] < y.p[axis]; });
    std::swap(items[b], items[mid]);
    uint32_t w = word[b];
    w = (w & ~3u) + (uint32_t)axis;
    split[b] = items[b].p[axis];
    bool has_left = b < mid;
    w = (w & ~4u) + (has_left ? 4u : 0u);
    uint32_t right = (mid + 1 < e) ? (uint32_t)(mid + 1) : kNoChild;
    w = (right << 3) + (w & 7u);
    word[b] = w;
    bool par = (e - b) > 65536 && has_left && mid + 1 < e && threads.fetch_add(1) < max_threads;
    if (par) {
      std::thread t([this, b, mid, depth] {
        build(b + 1, mid + 1, depth + 1);
        threads.fetch_sub(1);
      });
      build(mid + 1, e, depth + 1);
      t.join();
    } else {
      if ((e - b) > 65536 && has_left && mid + 1 < e) threads.fetch_sub(1);
      if (has_left) build(b + 1, mid + 1, depth + 1);
      if (mid + 1 < e) build(mid + 1, e, depth + 1);
    }
  }
};

}  // namespace

void build_photon_tree(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                       const uint16_t *normal, KdLayout &out) {
  KdBuilder kb;
  kb.max_threads = std::max(1, (int)std::thread::hardware_concurrency());
  kb.items.resize(n);
  for (uint64_t i = 0; i < n; ++i) kb.items[i] = KdItem{{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, (uint32_t)i};
  kb.word.assign(n, 0u);
  kb.split.assign(n, 0.0f);
  if (n > 0) kb.build(0, n, 0);
  out.nodes.resize(n);
  out.photons.resize(n);
  out.perm.resize(n);
  out.max_depth = kb.max_depth.load();
  for (uint64_t i = 0; i < n; ++i) {
    const KdItem &it = kb.items[i];
    DevKdNode &d = out.nodes[i];
    d.p[0] = it.p[0];
    d.p[1] = it.p[1];
    d.p[2] = it.p[2];
    d.word = kb.word[i];
    out.perm[i] = it.id;
    DevPhoton &ph = out.photons[i];
    for (int c = 0; c < 3; ++c) ph.w[c] = weight ? weight[3 * (size_t)it.id + c] : 0.0f;
    ph.codes = (uint32_t)(dir ? dir[it.id] : 0) | ((uint32_t)(normal ? normal[it.id] : 0) << 16);
  }
  out.psplit.assign(n, 0.0f);
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t w = out.nodes[i].word, axis = w & 3u;
    if (axis == 3u) continue;
    const float sv = out.nodes[i].p[axis];
    if (w & 4u) out.psplit[i + 1] = sv;
    if ((w >> 3) < n) out.psplit[w >> 3] = sv;
  }
}

const double *cdir_table() {
  static std::vector<double> table = [] {
    std::vector<double> t(3 * 65536);
    for (int i = 0; i < 65536; ++i) {
      unsigned char x = (unsigned char)((i & 0x1FC0) >> 6);
      unsigned char y = (unsigned char)(i & 0x3F);
      if (x + y >= 127) {
        x = (unsigned char)(127 - x);
        y = (unsigned char)(127 - y);
      }
      double v[3] = {x + 0.5, y + 0.5, (double)(126 - x - y)};
      double l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (l2 > 0.0) {
        double inv = 1.0 / std::sqrt(l2);
        for (int c = 0; c < 3; ++c) v[c] = v[c] * inv;
      }
      if (i & (1 << 13)) v[0] = -v[0];
      if (i & (1 << 14)) v[1] = -v[1];
      if (i & (1 << 15)) v[2] = -v[2];
      for (int c = 0; c < 3; ++c) t[3 * i + c] = v[c];
    }
    return t;
  }();
  return table.data();
}

}  // namespace skw
