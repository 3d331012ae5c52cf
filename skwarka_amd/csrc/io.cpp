// On-disk formats of the built structures, for checkpoint / resume without rebuilding: the scene
// (flattened ternary tree, triangles, instances, meshes and primitives), a photon map (kd-tree in
// tree order, payload, permutation, parent splits) and an irradiance photon map (its photons and
// lookup distance, and its kd-tree). The reference keeps these only in memory.
//
// A file is: 8-byte magic, uint32 format version, uint32 byte-order mark 0x01020304, then
// sections (uint32 tag, uint32 0, uint64 byte count, bytes), then the FNV-1a 64 hash of every
// preceding byte. Readers check the magic, version, byte order, every section's size against the
// counts it implies, and the hash; a file that fails any check is rejected with SKW_INVALID_ARG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace {

constexpr uint32_t kVersion = 1;
constexpr uint32_t kByteOrder = 0x01020304u;
const char kSceneMagic[8] = {'S', 'K', 'W', 'S', 'C', 'E', 'N', 'E'};
const char kPhotonMagic[8] = {'S', 'K', 'W', 'P', 'H', 'M', 'A', 'P'};
const char kIrrMagic[8] = {'S', 'K', 'W', 'I', 'R', 'M', 'A', 'P'};

struct Fnv {
  uint64_t h = 0xcbf29ce484222325ull;
  void add(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= 0x100000001b3ull;
    }
  }
};

class Writer {
 public:
  explicit Writer(const char *magic) {
    raw(magic, 8);
    pod(kVersion);
    pod(kByteOrder);
  }
  template <class T>
  void pod(const T &v) {
    raw(&v, sizeof(T));
  }
  template <class T>
  void section(uint32_t tag, const std::vector<T> &v) {
    section(tag, v.data(), v.size() * sizeof(T));
  }
  void section(uint32_t tag, const void *p, uint64_t bytes) {
    pod(tag);
    pod(uint32_t(0));
    pod(bytes);
    raw(p, bytes);
  }
  skw_status save(const char *path) {
    std::vector<unsigned char> out = buf_;
    const uint64_t h = hash_.h;
    out.insert(out.end(), reinterpret_cast<const unsigned char *>(&h), reinterpret_cast<const unsigned char *>(&h) + 8);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(SKW_INVALID_ARG, std::string("cannot open for writing: ") + path);
    const size_t n = std::fwrite(out.data(), 1, out.size(), f);
    const int c = std::fclose(f);
    if (n != out.size() || c != 0) return fail(SKW_INVALID_ARG, std::string("write failed: ") + path);
    return SKW_OK;
  }

 private:
  void raw(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    buf_.insert(buf_.end(), c, c + n);
    hash_.add(p, n);
  }
  std::vector<unsigned char> buf_;
  Fnv hash_;
};

class Reader {
 public:
  skw_status open(const char *path, const char *magic) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SKW_INVALID_ARG, std::string("cannot open: ") + path);
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size < 24) {
      std::fclose(f);
      return fail(SKW_INVALID_ARG, "file too short");
    }
    buf_.resize((size_t)size);
    const size_t n = std::fread(buf_.data(), 1, buf_.size(), f);
    std::fclose(f);
    if (n != buf_.size()) return fail(SKW_INVALID_ARG, "read failed");
    Fnv h;
    h.add(buf_.data(), buf_.size() - 8);
    uint64_t stored;
    std::memcpy(&stored, buf_.data() + buf_.size() - 8, 8);
    end_ = buf_.size() - 8;
    if (std::memcmp(buf_.data(), magic, 8) != 0) return fail(SKW_INVALID_ARG, "not a file of this kind");
    pos_ = 8;
    uint32_t version = 0, order = 0;
    if (!pod(version) || !pod(order)) return fail(SKW_INVALID_ARG, "truncated header");
    if (order != kByteOrder) return fail(SKW_INVALID_ARG, "byte order differs from this machine's");
    if (version != kVersion) return fail(SKW_INVALID_ARG, "unsupported format version");
    if (h.h != stored) return fail(SKW_INVALID_ARG, "checksum mismatch (corrupt file)");
    return SKW_OK;
  }
  template <class T>
  bool pod(T &v) {
    if (pos_ + sizeof(T) > end_) return false;
    std::memcpy(&v, buf_.data() + pos_, sizeof(T));
    pos_ += sizeof(T);
    return true;
  }
  // The next section, which must carry `tag` and hold `count` elements of T.
  template <class T>
  bool section(uint32_t tag, std::vector<T> &v, uint64_t count) {
    uint32_t t = 0, z = 0;
    uint64_t bytes = 0;
    if (!pod(t) || !pod(z) || !pod(bytes)) return false;
    if (t != tag || bytes != count * sizeof(T) || pos_ + bytes > end_) return false;
    v.resize(count);
    if (bytes) std::memcpy(v.data(), buf_.data() + pos_, bytes);
    pos_ += bytes;
    return true;
  }
  bool at_end() const { return pos_ == end_; }

 private:
  std::vector<unsigned char> buf_;
  size_t pos_ = 0, end_ = 0;
};

// ---------------------------------------------------------------- photon map sections

void write_kd(Writer &w, const KdLayout &h) {
  const uint64_t n = h.nodes.size();
  w.pod(n);
  w.pod(h.max_depth);
  w.section(1, h.nodes);
  w.section(2, h.photons);
  w.section(3, h.perm);
  w.section(4, h.psplit);
}

// The implicit kd-tree of a loaded photon map (KDTree.h:204-213): node i's left child is i + 1,
// its right child the word's index, both after i (which rules out cycles). Every node must be
// reached exactly once from the root and the stored depth must be the tree's: the lookups size
// their stacks by it.
bool check_kd(const KdLayout &h) {
  const uint64_t n = h.nodes.size();
  if (n == 0) return h.max_depth == 0;
  std::vector<uint32_t> depth(n, 0);
  std::vector<unsigned char> seen(n, 0);
  seen[0] = 1;
  uint32_t max_depth = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (!seen[i] || h.perm[i] >= n) return false;
    const uint32_t word = h.nodes[i].word, right = word >> 3;
    max_depth = std::max(max_depth, depth[i]);
    if ((word & 3u) == 3u) continue;  // a leaf: the lookups read none of its child fields
    if (word & 4u) {
      if (i + 1 >= n || seen[i + 1]) return false;
      seen[i + 1] = 1;
      depth[i + 1] = depth[i] + 1;
    }
    if (right != kNoChild) {
      if (right < i + ((word & 4u) ? 2 : 1) || right >= n || seen[right]) return false;
      seen[right] = 1;
      depth[right] = depth[i] + 1;
    }
  }
  return max_depth == h.max_depth;
}

bool read_kd(Reader &r, KdLayout &h) {
  uint64_t n = 0;
  if (!r.pod(n) || !r.pod(h.max_depth) || n >= (uint64_t)kNoChild) return false;
  if (!r.section(1, h.nodes, n) || !r.section(2, h.photons, n) || !r.section(3, h.perm, n) ||
      !r.section(4, h.psplit, n))
    return false;
  return check_kd(h);
}

// Depth of the subtree at `root` (root = 0), checking every index the traversal kernels follow:
// children blocks inside the node array, leaf triangle runs inside the triangle array, instance
// ranges inside the instance array (and only where allow_inst), each node reached once.
bool check_subtree(const TreeLayout &t, uint32_t root, bool allow_inst, uint32_t &depth_out) {
  const uint64_t nn = t.nodes.size(), nt = t.tris.size(), ni = t.insts.size();
  if (root >= nn) return false;
  std::vector<unsigned char> seen(nn, 0);
  std::vector<std::pair<uint32_t, uint32_t>> stack{{root, 0u}};
  seen[root] = 1;
  uint32_t md = 0;
  while (!stack.empty()) {
    const auto [i, d] = stack.back();
    stack.pop_back();
    md = std::max(md, d);
    const DevNode &nd = t.nodes[i];
    if ((nd.b & 3u) != 0) {
      const uint32_t mask = (nd.b >> 2) & 7u;
      const uint32_t nch = (uint32_t)__builtin_popcount(mask);
      if (mask == 0 || (uint64_t)nd.a + nch > nn) return false;
      for (uint32_t c = 0; c < nch; ++c) {
        const uint32_t k = nd.a + c;
        if (seen[k]) return false;
        seen[k] = 1;
        stack.push_back({k, d + 1});
      }
    } else {
      if ((uint64_t)nd.a + (nd.b >> 3) > nt) return false;
      if (nd.b & 4u) {
        if (!allow_inst || t.inst_range.size() != 2 * nn) return false;
        const uint32_t ib = t.inst_range[2 * (size_t)i], ie = t.inst_range[2 * (size_t)i + 1];
        if (ib > ie || ie > ni) return false;
      }
    }
  }
  depth_out = md;
  return true;
}

// True when every child's box lies inside its parent's box and no box coordinate is NaN, as in
// every tree the builders make (the traversal's fast paths rely on it; other trees take the
// reference's exact slab test at every node).
bool boxes_nest(const TreeLayout &t) {
  for (const DevNode &nd : t.nodes) {
    for (int k = 0; k < 3; ++k)
      if (!(nd.bmin[k] == nd.bmin[k]) || !(nd.bmax[k] == nd.bmax[k])) return false;
    if ((nd.b & 3u) == 0) continue;
    const uint32_t nch = (uint32_t)__builtin_popcount((nd.b >> 2) & 7u);
    for (uint32_t c = 0; c < nch; ++c) {
      const DevNode &ch = t.nodes[nd.a + c];
      for (int k = 0; k < 3; ++k)
        if (!(ch.bmin[k] >= nd.bmin[k]) || !(ch.bmax[k] <= nd.bmax[k])) return false;
    }
  }
  return true;
}

// Recomputes the depths the builder records and rejects a tree that would send a traversal out of
// its arrays or past its stack (the kernels trust these values).
bool check_tree(const TreeLayout &t) {
  if (t.nodes.empty()) return t.tris.empty() && t.insts.empty();
  uint32_t top = 0;
  if (!check_subtree(t, t.root, true, top)) return false;
  uint32_t inst_depth = 0;
  std::vector<uint32_t> roots;
  for (const DevInst &in : t.insts) {
    if (in.prim >= t.prims.size()) return false;
    roots.push_back(in.root);
  }
  std::sort(roots.begin(), roots.end());
  roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
  for (uint32_t r : roots) {
    uint32_t d = 0;
    if (!check_subtree(t, r, false, d)) return false;
    inst_depth = std::max(inst_depth, d);
  }
  const uint32_t need = top + 1 + (t.insts.empty() ? 0 : inst_depth + 1);
  return top == t.max_depth && need == t.stack_need;
}

}  // namespace

extern "C" {

skw_status skw_scene_save(const skw_scene *scene, const char *path) {
  if (!scene || !path) return fail(SKW_INVALID_ARG, "scene and path must not be NULL");
  const TreeLayout *tp = nullptr;
  skw_status hs = scene_host(scene, &tp);
  if (hs != SKW_OK) return hs;
  const TreeLayout &t = *tp;
  Writer w(kSceneMagic);
  w.pod(t.root);
  w.pod(t.max_depth);
  w.pod(t.stack_need);
  w.pod(t.leaves);
  for (int i = 0; i < 6; ++i) w.pod(t.bounds[i]);
  w.pod((uint64_t)t.nodes.size());
  w.pod((uint64_t)t.tris.size());
  w.pod((uint64_t)t.insts.size());
  w.pod((uint64_t)t.inst_range.size());
  w.pod((uint64_t)t.meshes.size());
  w.pod((uint64_t)t.prims.size());
  w.section(1, t.nodes);
  w.section(2, t.tris);
  w.section(3, t.insts);
  w.section(4, t.inst_range);
  for (const HostMesh &m : t.meshes) {
    w.pod((uint64_t)m.verts.size());
    w.pod((uint64_t)m.normals.size());
    w.pod((uint64_t)m.tris.size());
    w.pod((uint32_t)(m.invert_normals ? 1u : 0u));
    w.section(5, m.verts);
    w.section(6, m.normals);
    w.section(7, m.tris);
  }
  std::vector<double> xf;
  std::vector<uint32_t> meta;
  for (const HostPrim &p : t.prims) {
    xf.insert(xf.end(), p.m2w, p.m2w + 16);
    xf.insert(xf.end(), p.w2m, p.w2m + 16);
    meta.push_back(p.mesh);
    meta.push_back(p.identity ? 1u : 0u);
  }
  w.section(8, xf);
  w.section(9, meta);
  return w.save(path);
}

skw_status skw_scene_load(const char *path, uint32_t flags, skw_scene **out) {
  if (!path || !out) return fail(SKW_INVALID_ARG, "path and out must not be NULL");
  *out = nullptr;
  const bool host_only = (flags & SKW_SCENE_HOST_ONLY) != 0;
  if (!host_only && device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  Reader r;
  skw_status st = r.open(path, kSceneMagic);
  if (st != SKW_OK) return st;
  std::unique_ptr<skw_scene> s(new skw_scene());
  TreeLayout &t = s->host;
  uint64_t nn = 0, nt = 0, ni = 0, nr = 0, nm = 0, np = 0;
  bool ok = r.pod(t.root) && r.pod(t.max_depth) && r.pod(t.stack_need) && r.pod(t.leaves);
  for (int i = 0; i < 6 && ok; ++i) ok = r.pod(t.bounds[i]);
  ok = ok && r.pod(nn) && r.pod(nt) && r.pod(ni) && r.pod(nr) && r.pod(nm) && r.pod(np);
  ok = ok && r.section(1, t.nodes, nn) && r.section(2, t.tris, nt) && r.section(3, t.insts, ni) &&
       r.section(4, t.inst_range, nr);
  if (ok) t.meshes.resize(nm);
  for (uint64_t i = 0; i < nm && ok; ++i) {
    HostMesh &m = t.meshes[i];
    uint64_t v = 0, n = 0, tr = 0;
    uint32_t inv = 0;
    ok = r.pod(v) && r.pod(n) && r.pod(tr) && r.pod(inv) && r.section(5, m.verts, v) && r.section(6, m.normals, n) &&
         r.section(7, m.tris, tr);
    m.invert_normals = inv != 0;
    ok = ok && v % 3 == 0 && tr % 3 == 0 && (n == 0 || n == v);
    for (uint64_t k = 0; k < tr && ok; ++k) ok = m.tris[k] < v / 3;
  }
  std::vector<double> xf;
  std::vector<uint32_t> meta;
  ok = ok && r.section(8, xf, np * 32) && r.section(9, meta, np * 2) && r.at_end();
  if (ok) {
    t.prims.resize(np);
    for (uint64_t i = 0; i < np && ok; ++i) {
      HostPrim &p = t.prims[i];
      std::memcpy(p.m2w, xf.data() + 32 * i, 16 * sizeof(double));
      std::memcpy(p.w2m, xf.data() + 32 * i + 16, 16 * sizeof(double));
      p.mesh = meta[2 * i];
      p.identity = meta[2 * i + 1] != 0;
      ok = p.mesh < nm;
    }
  }
  ok = ok && (nn == 0 || t.root < nn);
  for (uint64_t i = 0; i < nt && ok; ++i) ok = t.tris[i].prim < np || t.tris[i].prim == kMiss;  // kMiss: instanced
  if (!ok) return fail(SKW_INVALID_ARG, "scene file: inconsistent sections");
  if (!check_tree(t))
    return fail(SKW_INVALID_ARG, "scene file: the tree's indices or recorded depths are inconsistent");
  t.nested = boxes_nest(t);
  st = finish_scene(s, host_only);
  if (st != SKW_OK) return st;
  *out = s.release();
  return SKW_OK;
}

skw_status skw_photon_map_save(const skw_photon_map *map, const char *path) {
  if (!map || !path) return fail(SKW_INVALID_ARG, "map and path must not be NULL");
  const KdLayout &h = photon_map_host(map);
  if (h.nodes.size() != map->n) return fail(SKW_HIP_ERROR, "photon tree download failed");
  Writer w(kPhotonMagic);
  write_kd(w, h);
  return w.save(path);
}

skw_status skw_photon_map_load(const char *path, uint32_t flags, skw_photon_map **out) {
  if (!path || !out) return fail(SKW_INVALID_ARG, "path and out must not be NULL");
  *out = nullptr;
  const bool host_only = (flags & SKW_SCENE_HOST_ONLY) != 0;
  if (!host_only && device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  Reader r;
  skw_status st = r.open(path, kPhotonMagic);
  if (st != SKW_OK) return st;
  std::unique_ptr<skw_photon_map> m(new skw_photon_map());
  if (!read_kd(r, m->host) || !r.at_end()) return fail(SKW_INVALID_ARG, "photon map file: inconsistent sections");
  if (host_only) {
    m->n = m->host.nodes.size();
    m->max_depth = m->host.max_depth;
  } else if ((st = finish_photon_map(m.get())) != SKW_OK) {
    return st;
  }
  *out = m.release();
  return SKW_OK;
}

skw_status skw_irradiance_map_save(const skw_irradiance_map *map, const char *path) {
  if (!map || !path) return fail(SKW_INVALID_ARG, "map and path must not be NULL");
  Writer w(kIrrMagic);
  const uint64_t n = map->normal.size();
  w.pod(n);
  w.pod(map->max_dist);
  w.section(1, map->pos);
  w.section(2, map->normal);
  w.section(3, map->e_ext);
  w.section(4, map->e_int);
  const uint32_t has_tree = map->map ? 1u : 0u;
  w.pod(has_tree);
  if (map->map) {
    const KdLayout &h = photon_map_host(map->map);
    if (h.nodes.size() != map->map->n) return fail(SKW_HIP_ERROR, "photon tree download failed");
    write_kd(w, h);
  }
  return w.save(path);
}

skw_status skw_irradiance_map_load(const char *path, skw_irradiance_map **out) {
  if (!path || !out) return fail(SKW_INVALID_ARG, "path and out must not be NULL");
  *out = nullptr;
  if (device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  Reader r;
  skw_status st = r.open(path, kIrrMagic);
  if (st != SKW_OK) return st;
  std::unique_ptr<skw_irradiance_map> m(new skw_irradiance_map());
  uint64_t n = 0;
  uint32_t has_tree = 0;
  bool ok = r.pod(n) && r.pod(m->max_dist) && r.section(1, m->pos, 3 * n) && r.section(2, m->normal, n) &&
            r.section(3, m->e_ext, 3 * n) && r.section(4, m->e_int, 3 * n) && r.pod(has_tree);
  ok = ok && (has_tree == (n > 0 ? 1u : 0u));
  if (ok && has_tree) {
    std::unique_ptr<skw_photon_map> pm(new skw_photon_map());
    ok = read_kd(r, pm->host) && pm->host.nodes.size() == n;
    if (ok) {
      if ((st = finish_photon_map(pm.get())) != SKW_OK) return st;
      if ((st = photon_map_nbox(pm.get(), nullptr)) != SKW_OK) return st;
      m->map = pm.release();
    }
  }
  if (!ok || !r.at_end()) return fail(SKW_INVALID_ARG, "irradiance map file: inconsistent sections");
  if (n > 0 && (st = upload_irradiance(m.get())) != SKW_OK) return st;
  *out = m.release();
  return SKW_OK;
}

}  // extern "C"
