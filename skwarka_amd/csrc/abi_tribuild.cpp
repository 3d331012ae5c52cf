// Triangle tree built on the device (skw_scene_create with SKW_SCENE_DEVICE_BUILD), node for node
// the tree of the host builder (TriangleAccelerator.cpp:22-147, 505-729):
//   1. the host reads and cleans the meshes as the host builder does and uploads the scene's
//      primitive and mesh tables (the same tables every scene uploads);
//   2. k_tb_prep moves the triangles to world space in the builder's order with the builder's
//      arithmetic and forms their float boxes;
//   3. the kernels of tri_build.hip / tri_build_wide.hip build the tree level by level (node
//      boxes, split search, the in-place three-way partition);
//   4. the k_tl_* kernels lay the nodes out depth first as the host builder's place() does and
//      gather the triangle array in its final order, straight into the scene's device arrays.
// The host copy of the tree (for export and save) is downloaded on first use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace {

constexpr uint32_t kNoChild = 0xFFFFFFFFu;

// Device buffers of one build, released on every exit path.
struct Buffers {
  hipStream_t st = nullptr;
  std::vector<void *> held;
  ~Buffers() {
    if (st) (void)hipStreamSynchronize(st);
    for (void *p : held) pool_free(p, st);
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
  template <class T>
  hipError_t alloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = pool_alloc((void **)p, std::max<size_t>(1, count) * sizeof(T), st);
    if (e == hipSuccess) held.push_back((void *)*p);
    return e;
  }
};

struct BuildConfig {
  uint32_t small_max;  // nodes of at most this many triangles: one wave builds the subtree
  uint32_t wide_min;   // nodes of more than this many: spread over chunks of `chunk` triangles
  uint32_t chunk;
};

uint32_t ceil_log2(uint32_t v) {
  uint32_t r = 0;
  while (r < 32 && (1ull << r) < v) ++r;
  return r;
}

// One build over the boxes and order that k_tb_prep leaves in a.box / a.ref, with `cap` node
// records (d_nodes_out, ids as allocated; count set); overflow when a subtree build ran out of
// records.
skw_status build_levels(Buffers &B, TriWideArgs &wa, uint32_t n, const BuildConfig &cfg, uint32_t cap,
                        TriBuildNode **d_nodes_out, uint32_t &count, bool &overflow) {
  TriBuildArgs &a = wa.b;
  uint32_t *d_counts = nullptr;  // node count, next big, next small, overflow, next wide
  SKW_HIP(B.alloc(&d_counts, 8), "tree build allocation");
  const uint32_t wcap = n / cfg.wide_min + 2;
  TriBuildNode *d_nodes = nullptr;
  uint32_t *cur[3] = {nullptr, nullptr, nullptr}, *nxt[3] = {nullptr, nullptr, nullptr};  // wide, big, small
  SKW_HIP(B.alloc(&d_nodes, cap), "tree build allocation");
  for (int l = 0; l < 3; ++l) {
    SKW_HIP(B.alloc(&cur[l], cap), "tree build allocation");
    SKW_HIP(B.alloc(&nxt[l], cap), "tree build allocation");
  }
  TriBuildNode root;
  std::memset(&root, 0, sizeof(root));
  root.te = n;
  root.first_child = kNoChild;
  const int root_list = n > cfg.wide_min ? 0 : (n > cfg.small_max ? 1 : 2);
  uint32_t counts[3] = {0, 0, 0};
  counts[root_list] = 1;
  uint32_t h_counts[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t zero = 0;
  SKW_HIP(hipMemcpyAsync(d_nodes, &root, sizeof(root), hipMemcpyHostToDevice, B.st), "upload root");
  SKW_HIP(hipMemcpyAsync(cur[root_list], &zero, sizeof(zero), hipMemcpyHostToDevice, B.st), "upload root list");
  count = 1;
  overflow = false;
  a.nodes = d_nodes;
  a.node_count = d_counts;
  a.next_big_n = d_counts + 1;
  a.next_small_n = d_counts + 2;
  a.overflow = d_counts + 3;
  a.next_wide_n = d_counts + 4;
  a.node_cap = cap;
  a.small_max = cfg.small_max;
  a.wide_min = cfg.wide_min;
  while (counts[0] + counts[1] + counts[2] > 0) {
    const uint64_t need = (uint64_t)count + 3ull * (counts[0] + counts[1]);
    if (need > cap || counts[0] > wcap) {
      // the level's children would not fit: a larger capacity is the caller's retry
      overflow = true;
      return SKW_OK;
    }
    h_counts[0] = count;
    for (int i = 1; i < 8; ++i) h_counts[i] = 0;
    SKW_HIP(hipMemcpyAsync(d_counts, h_counts, 8 * sizeof(uint32_t), hipMemcpyHostToDevice, B.st), "set counts");
    a.next_wide = nxt[0];
    a.next_big = nxt[1];
    a.next_small = nxt[2];
    if (counts[0] > 0) {
      a.list = cur[0];
      a.list_n = counts[0];
      SKW_HIP(launch_tri_wide_setup(wa, B.st), "tree build wide setup");
      uint32_t info[2] = {0, 0};
      SKW_HIP(hipMemcpyAsync(info, wa.info_out, sizeof(info), hipMemcpyDeviceToHost, B.st), "read chunks");
      SKW_HIP(hipStreamSynchronize(B.st), "tree build wide setup");
      wa.n_chunks = info[0];
      SKW_HIP(launch_tri_wide_level(wa, ceil_log2(info[1]) + 1, B.st), "tree build wide level");
    }
    a.list = cur[1];
    a.list_n = counts[1];
    SKW_HIP(launch_tri_build_level(a, B.st), "tree build level");
    a.list = cur[2];
    a.list_n = counts[2];
    SKW_HIP(launch_tri_build_small(a, B.st), "tree build subtrees");
    SKW_HIP(hipMemcpyAsync(h_counts, d_counts, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, B.st), "read counts");
    SKW_HIP(hipStreamSynchronize(B.st), "tree build level");
    if (h_counts[3] != 0 || h_counts[0] > cap) {
      overflow = true;
      return SKW_OK;
    }
    count = h_counts[0];
    counts[0] = h_counts[4];
    counts[1] = h_counts[1];
    counts[2] = h_counts[2];
    for (int l = 0; l < 3; ++l) std::swap(cur[l], nxt[l]);
  }
  *d_nodes_out = d_nodes;
  return SKW_OK;
}

// The depth-first layout of the built records into the scene's node and triangle arrays.
skw_status layout_on_device(Buffers &B, const TriBuildNode *d_hn, uint32_t count, uint32_t n, const float *d_tv,
                            const uint32_t *d_tprim, const uint32_t *d_ttri, const uint32_t *d_ref, skw_scene *s) {
  TriLayoutArgs L;
  std::memset(&L, 0, sizeof(L));
  L.hn = d_hn;
  L.count = count;
  L.n = n;
  L.tv = d_tv;
  L.tprim = d_tprim;
  L.ttri = d_ttri;
  L.ref = d_ref;
  // hist, fill (kTriLayoutDepths each), flags, leaves: one block, read back at once
  uint32_t *d_small = nullptr, *d_off = nullptr;
  SKW_HIP(B.alloc(&d_small, 2 * kTriLayoutDepths + 2), "layout allocation");
  SKW_HIP(B.alloc(&d_off, kTriLayoutDepths + 1), "layout allocation");
  L.hist = d_small;
  L.fill = d_small + kTriLayoutDepths;
  L.flags = d_small + 2 * kTriLayoutDepths;
  L.leaves = d_small + 2 * kTriLayoutDepths + 1;
  L.level_off = d_off;
  SKW_HIP(B.alloc(&L.by_depth, count), "layout allocation");
  SKW_HIP(B.alloc(&L.size, count), "layout allocation");
  SKW_HIP(B.alloc(&L.cval, count), "layout allocation");
  SKW_HIP(B.alloc(&L.idx, count), "layout allocation");
  SKW_HIP(hipMalloc((void **)&s->d_nodes, (size_t)count * sizeof(DevNode)), "scene nodes");
  SKW_HIP(hipMalloc((void **)&s->d_tris, (size_t)n * sizeof(DevTri)), "scene triangles");
  L.nodes = s->d_nodes;
  L.tris = s->d_tris;
  SKW_HIP(hipMemsetAsync(d_small, 0, (2 * kTriLayoutDepths + 2) * sizeof(uint32_t), B.st), "layout clear");
  SKW_HIP(launch_tri_layout_hist(L, B.st), "layout histogram");
  std::vector<uint32_t> small(2 * kTriLayoutDepths + 2);
  SKW_HIP(hipMemcpyAsync(small.data(), d_small, small.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, B.st),
          "layout histogram");
  SKW_HIP(hipStreamSynchronize(B.st), "layout histogram");
  if (small[2 * kTriLayoutDepths] != 0) return fail(SKW_HIP_ERROR, "device tree build: a node record is out of range");
  std::vector<uint32_t> off(kTriLayoutDepths + 1, 0);
  uint32_t max_depth = 0;
  for (uint32_t d = 0; d < kTriLayoutDepths; ++d) {
    off[d + 1] = off[d] + small[d];
    if (small[d]) max_depth = d;
  }
  if (off[kTriLayoutDepths] != count || small[0] != 1)
    return fail(SKW_HIP_ERROR, "device tree build: the records do not form one tree");
  SKW_HIP(hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, B.st),
          "layout offsets");
  SKW_HIP(launch_tri_layout_scatter(L, B.st), "layout scatter");
  for (uint32_t d = max_depth + 1; d-- > 0;) {
    L.level_begin = off[d];
    L.level_end = off[d + 1];
    SKW_HIP(launch_tri_layout_size(L, B.st), "layout sizes");
  }
  for (uint32_t d = 0; d <= max_depth; ++d) {
    L.level_begin = off[d];
    L.level_end = off[d + 1];
    SKW_HIP(launch_tri_layout_place(L, B.st), "layout placement");
  }
  SKW_HIP(launch_tri_layout_tris(L, B.st), "layout triangles");
  TriBuildNode root;
  SKW_HIP(hipMemcpyAsync(&root, d_hn, sizeof(root), hipMemcpyDeviceToHost, B.st), "root record");
  SKW_HIP(hipStreamSynchronize(B.st), "layout");
  TreeLayout &t = s->host;
  t.root = 0;
  t.max_depth = max_depth;
  t.stack_need = max_depth + 1;
  t.leaves = small[2 * kTriLayoutDepths + 1];
  for (int i = 0; i < 3; ++i) {
    t.bounds[i] = root.bmin[i];
    t.bounds[3 + i] = root.bmax[i];
  }
  s->n_nodes = count;
  s->n_tris = n;
  s->host_tree = false;
  return SKW_OK;
}

}  // namespace

namespace skw {

skw_status build_triangle_tree_device(const skw_scene_desc &desc, skw_scene *s, bool &built) {
  built = false;
  std::string err;
  std::vector<uint64_t> goff;
  std::vector<uint32_t> user;
  bool instanced = false;
  if (!read_plain_scene(desc, s->host, goff, user, instanced, err)) return fail(SKW_INVALID_ARG, err);
  if (instanced)
    return fail(SKW_UNSUPPORTED, "the device tree build takes scenes without instanced meshes (a mesh used by "
                                 "several primitives); build such scenes on the host");
  const uint64_t n64 = goff.back();
  if (n64 == 0) return SKW_OK;  // the empty tree: the caller builds it on the host
  const uint32_t n = (uint32_t)n64;
  BuildConfig cfg;
  cfg.small_max = (uint32_t)std::min(64, std::max(1, env_int("SKW_TRIBUILD_SMALL_MAX", 64)));
  cfg.wide_min = (uint32_t)std::max<int>((int)std::max<uint32_t>(cfg.small_max, 256),
                                         env_int("SKW_TRIBUILD_WIDE_MIN", 16384));
  cfg.chunk = 256u * (uint32_t)std::min(64, std::max(1, env_int("SKW_TRIBUILD_CHUNK", 4096) / 256));
  skw_status st = scene_tables(s, false);
  if (st != SKW_OK) return st;

  Buffers B;
  SKW_HIP(hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking), "hipStreamCreate");
  TriPrepArgs p;
  std::memset(&p, 0, sizeof(p));
  p.meshes = s->d_meshes;
  p.verts = s->d_mesh_verts;
  p.tris = s->d_mesh_tris;
  p.prims = s->d_prims;
  p.n_meshes = (uint32_t)s->host.meshes.size();
  p.n = n;
  uint64_t *d_goff = nullptr;
  uint32_t *d_user = nullptr;
  SKW_HIP(B.alloc(&d_goff, goff.size()), "tree build allocation");
  SKW_HIP(B.alloc(&d_user, user.size()), "tree build allocation");
  SKW_HIP(hipMemcpyAsync(d_goff, goff.data(), goff.size() * sizeof(uint64_t), hipMemcpyHostToDevice, B.st),
          "upload mesh offsets");
  SKW_HIP(hipMemcpyAsync(d_user, user.data(), user.size() * sizeof(uint32_t), hipMemcpyHostToDevice, B.st),
          "upload mesh primitives");
  p.goff = d_goff;
  p.user = d_user;
  SKW_HIP(B.alloc(&p.tv, 9 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&p.box, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&p.tprim, n), "tree build allocation");
  SKW_HIP(B.alloc(&p.ttri, n), "tree build allocation");
  SKW_HIP(B.alloc(&p.ref, n), "tree build allocation");
  SKW_HIP(B.alloc(&p.flags, 4), "tree build allocation");

  TriWideArgs wa;
  std::memset(&wa, 0, sizeof(wa));
  TriBuildArgs &a = wa.b;
  a.box = p.box;
  a.ref = p.ref;
  SKW_HIP(B.alloc(&a.box2, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&a.ref2, n), "tree build allocation");
  uint32_t **scratch[9] = {&a.s_cls, &a.s_rf, &a.s_nb, &a.s_G, &a.s_Rpos, &a.s_order, &a.s_elem, &a.s_ptr, &a.s_dest};
  for (uint32_t **sc : scratch) SKW_HIP(B.alloc(sc, n), "tree build allocation");
  const uint32_t wcap = n / cfg.wide_min + 2;
  wa.chunk_cap = n / cfg.chunk + wcap + 2;
  wa.chunk = cfg.chunk;
  SKW_HIP(B.alloc(&wa.w, wcap), "tree build allocation");
  unsigned char *d_bins = nullptr;
  SKW_HIP(B.alloc(&d_bins, 3 * (size_t)wcap * tri_build_bins_bytes()), "tree build allocation");
  wa.bins = d_bins;
  SKW_HIP(B.alloc(&wa.chunk_w, wa.chunk_cap), "tree build allocation");
  SKW_HIP(B.alloc(&wa.ccnt, 8 * (size_t)wa.chunk_cap), "tree build allocation");
  SKW_HIP(B.alloc(&wa.cbase, 8 * (size_t)wa.chunk_cap), "tree build allocation");
  SKW_HIP(B.alloc(&wa.info_out, 2), "tree build allocation");

  // node records: the level kernel's children are counted before each level; subtree builds check
  // the capacity themselves and the build restarts (from the prepared triangles) with more records
  // when they run out
  uint32_t cap = std::max<uint32_t>(1024, n + n / 2 + 64);
  for (int attempt = 0;; ++attempt) {
    uint32_t flags = 0;
    SKW_HIP(hipMemsetAsync(p.flags, 0, 4 * sizeof(uint32_t), B.st), "tree build clear");
    SKW_HIP(launch_tri_prep(p, B.st), "world-space triangles");
    SKW_HIP(hipMemcpyAsync(&flags, p.flags, sizeof(flags), hipMemcpyDeviceToHost, B.st), "read flags");
    SKW_HIP(hipStreamSynchronize(B.st), "world-space triangles");
    if (flags & 1u) return fail(SKW_UNSUPPORTED, "the device tree build needs finite vertices");
    TriBuildNode *d_hn = nullptr;
    uint32_t count = 0;
    bool overflow = false;
    st = build_levels(B, wa, n, cfg, cap, &d_hn, count, overflow);
    if (st != SKW_OK) return st;
    if (overflow) {
      if (attempt >= 3 || cap >= 0x40000000u) return fail(SKW_OUT_OF_MEMORY, "device tree build: too many nodes");
      cap *= 4;
      continue;
    }
    st = layout_on_device(B, d_hn, count, n, p.tv, p.tprim, p.ttri, a.ref, s);
    if (st != SKW_OK) return st;
    built = true;
    return SKW_OK;
  }
}

skw_status scene_host(const skw_scene *scene, const TreeLayout **out) {
  skw_scene *s = const_cast<skw_scene *>(scene);
  std::lock_guard<std::mutex> lock(s->host_mu);
  if (!s->host_tree) {
    int cur = 0;
    SKW_HIP(hipGetDevice(&cur), "hipGetDevice");
    SKW_HIP(hipSetDevice(s->device), "hipSetDevice");
    std::vector<DevNode> nodes(s->n_nodes);
    std::vector<DevTri> tris(s->n_tris);
    hipError_t e = hipMemcpy(nodes.data(), s->d_nodes, nodes.size() * sizeof(DevNode), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(tris.data(), s->d_tris, tris.size() * sizeof(DevTri), hipMemcpyDeviceToHost);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "download tree");
    s->host.nodes.swap(nodes);
    s->host.tris.swap(tris);
    s->host_tree = true;
  }
  *out = &s->host;
  return SKW_OK;
}

}  // namespace skw
