// Triangle tree built on the device (skw_scene_create with SKW_SCENE_DEVICE_BUILD): the host
// prepares the world-space triangles and their boxes exactly as the host builder does, the
// kernels of tri_build.hip build the tree level by level (node boxes, split search, the in-place
// three-way partition), and the host lays the nodes out depth first as the host builder's
// place() does, so the scene is node for node the one the host builder makes
// (TriangleAccelerator.cpp:22-147, 505-729).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
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
  template <class T>
  void release(T *p) {
    auto it = std::find(held.begin(), held.end(), (void *)p);
    if (it != held.end()) {
      pool_free(p, st);
      held.erase(it);
    }
  }
};

// Lays the nodes out as the host builder's place() does: node by node depth first (children left,
// middle, right), each internal node's children as one block appended when the node is placed.
void place_all(const std::vector<TriBuildNode> &hn, TreeLayout &out) {
  out.nodes.clear();
  out.nodes.reserve(hn.size());
  out.nodes.resize(1);
  out.leaves = 0;
  std::vector<std::pair<uint32_t, uint32_t>> st;  // (node id, layout index)
  st.reserve(1024);
  st.push_back({0u, 0u});
  while (!st.empty()) {
    const uint32_t id = st.back().first, idx = st.back().second;
    st.pop_back();
    const TriBuildNode &n = hn[id];
    DevNode &d = out.nodes[idx];
    for (int i = 0; i < 3; ++i) {
      d.bmin[i] = n.bmin[i];
      d.bmax[i] = n.bmax[i];
    }
    if (n.first_child == kNoChild) {
      d.a = n.tb;
      d.b = (n.te - n.tb) << 3;
      out.leaves++;
      continue;
    }
    const uint32_t flags = n.info & 3u, mask = (n.info >> 2) & 7u, nch = (uint32_t)__builtin_popcount(mask);
    const uint32_t first = (uint32_t)out.nodes.size();
    d.a = first;
    d.b = flags | (mask << 2);
    out.nodes.resize(out.nodes.size() + nch);
    for (uint32_t k = nch; k-- > 0;) st.push_back({n.first_child + k, first + k});
  }
}

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

// One build on the device with `cap` node records: the nodes (ids as allocated) and the final
// triangle order; overflow when a subtree build ran out of records.
skw_status build_levels(const PlainTriangles &pt, uint32_t n, const BuildConfig &cfg,
                        const std::vector<uint32_t> &iota, uint32_t cap, std::vector<TriBuildNode> &hn,
                        std::vector<uint32_t> &order, bool &overflow) {
  Buffers B;
  SKW_HIP(hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking), "hipStreamCreate");
  TriWideArgs wa;
  std::memset(&wa, 0, sizeof(wa));
  TriBuildArgs &a = wa.b;
  SKW_HIP(B.alloc(&a.box, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&a.box2, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&a.ref, n), "tree build allocation");
  SKW_HIP(B.alloc(&a.ref2, n), "tree build allocation");
  uint32_t **scratch[9] = {&a.s_cls, &a.s_rf, &a.s_nb, &a.s_G, &a.s_Rpos, &a.s_order, &a.s_elem, &a.s_ptr, &a.s_dest};
  for (uint32_t **s : scratch) SKW_HIP(B.alloc(s, n), "tree build allocation");
  // counts: node count, next big, next small, overflow, next wide
  uint32_t *d_counts = nullptr;
  SKW_HIP(B.alloc(&d_counts, 8), "tree build allocation");
  // wide-node state: at most n / wide_min nodes per level, their chunks
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
  SKW_HIP(hipMemcpyAsync(a.box, pt.box.data(), pt.box.size() * sizeof(float), hipMemcpyHostToDevice, B.st),
          "upload boxes");
  SKW_HIP(hipMemcpyAsync(a.ref, iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, B.st), "upload ids");
  TriBuildNode *d_nodes = nullptr;
  // current / next lists: wide, big, small
  uint32_t *cur[3] = {nullptr, nullptr, nullptr}, *nxt[3] = {nullptr, nullptr, nullptr};
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
  uint32_t counts[3] = {0, 0, 0};  // wide, big, small
  counts[root_list] = 1;
  uint32_t h_counts[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t zero = 0;
  SKW_HIP(hipMemcpyAsync(d_nodes, &root, sizeof(root), hipMemcpyHostToDevice, B.st), "upload root");
  SKW_HIP(hipMemcpyAsync(cur[root_list], &zero, sizeof(zero), hipMemcpyHostToDevice, B.st), "upload root list");
  uint32_t count = 1;
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
      break;
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
      break;
    }
    count = h_counts[0];
    counts[0] = h_counts[4];
    counts[1] = h_counts[1];
    counts[2] = h_counts[2];
    for (int l = 0; l < 3; ++l) std::swap(cur[l], nxt[l]);
  }
  if (overflow) return SKW_OK;
  hn.resize(count);
  order.resize(n);
  SKW_HIP(hipMemcpyAsync(hn.data(), d_nodes, (size_t)count * sizeof(TriBuildNode), hipMemcpyDeviceToHost, B.st),
          "download nodes");
  SKW_HIP(hipMemcpyAsync(order.data(), a.ref, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, B.st),
          "download triangle order");
  SKW_HIP(hipStreamSynchronize(B.st), "download tree");
  return SKW_OK;
}

// The host layout of a device-built tree: depth-first placement, depths, bounds, triangles.
skw_status layout(const PlainTriangles &pt, uint32_t n, const std::vector<TriBuildNode> &hn,
                  const std::vector<uint32_t> &order, TreeLayout &out) {
  // children are allocated after their parent: ids grow along every path, so the placement ends
  for (size_t i = 0; i < hn.size(); ++i) {
    const TriBuildNode &nd = hn[i];
    if (nd.first_child == kNoChild) continue;
    const uint64_t nch = (uint64_t)__builtin_popcount((nd.info >> 2) & 7u);
    if (nd.first_child <= i || nch == 0 || nd.first_child + nch > hn.size())
      return fail(SKW_HIP_ERROR, "device tree build: child index out of range");
  }
  place_all(hn, out);
  out.root = 0;
  uint32_t depth = 0;
  for (const TriBuildNode &nd : hn) depth = std::max(depth, nd.info >> 16);
  out.max_depth = depth;
  out.stack_need = depth + 1;
  for (int i = 0; i < 3; ++i) {
    out.bounds[i] = hn[0].bmin[i];
    out.bounds[3 + i] = hn[0].bmax[i];
  }
  out.tris.resize(n);
  std::atomic<bool> bad{false};
  parallel_ranges(n, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t pos = lo; pos < hi; ++pos) {
      const uint32_t id = order[pos];
      if (id >= n) {
        bad = true;
        return;
      }
      DevTri &t = out.tris[pos];
      std::memcpy(t.v0, &pt.tv[9 * (size_t)id], 3 * sizeof(float));
      std::memcpy(t.v1, &pt.tv[9 * (size_t)id + 3], 3 * sizeof(float));
      std::memcpy(t.v2, &pt.tv[9 * (size_t)id + 6], 3 * sizeof(float));
      t.prim = pt.tprim[id];
      t.tri = pt.ttri[id];
      t.pad = 0;
    }
  });
  if (bad) return fail(SKW_HIP_ERROR, "device tree build: triangle order out of range");
  return SKW_OK;
}

}  // namespace

namespace skw {

skw_status build_triangle_tree_device(const skw_scene_desc &desc, TreeLayout &out) {
  std::string err;
  PlainTriangles pt;
  bool instanced = false;
  if (!prepare_plain_triangles(desc, out, pt, instanced, err)) return fail(SKW_INVALID_ARG, err);
  if (instanced)
    return fail(SKW_UNSUPPORTED, "the device tree build takes scenes without instanced meshes (a mesh used by "
                                 "several primitives); build such scenes on the host");
  const uint64_t n64 = pt.tprim.size();
  if (n64 == 0) {
    // the empty tree: one leaf whose box stays (+inf, -inf)
    if (!build_triangle_tree(desc, out, err)) return fail(SKW_INVALID_ARG, err);
    return SKW_OK;
  }
  if (!pt.finite_boxes) return fail(SKW_UNSUPPORTED, "the device tree build needs finite vertices");
  const uint32_t n = (uint32_t)n64;
  BuildConfig cfg;
  cfg.small_max = (uint32_t)std::min(64, std::max(1, env_int("SKW_TRIBUILD_SMALL_MAX", 64)));
  cfg.wide_min = (uint32_t)std::max<int>((int)std::max<uint32_t>(cfg.small_max, 256),
                                         env_int("SKW_TRIBUILD_WIDE_MIN", 16384));
  cfg.chunk = 256u * (uint32_t)std::min(64, std::max(1, env_int("SKW_TRIBUILD_CHUNK", 4096) / 256));
  std::vector<uint32_t> iota(n);
  parallel_ranges(n, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) iota[i] = (uint32_t)i;
  });
  // node records: the level kernel's children are counted before each level; subtree builds check
  // the capacity themselves and the build restarts with more records when they run out
  uint32_t cap = std::max<uint32_t>(1024, n + n / 2 + 64);
  for (int attempt = 0;; ++attempt) {
    std::vector<TriBuildNode> hn;
    std::vector<uint32_t> order;
    bool overflow = false;
    skw_status st = build_levels(pt, n, cfg, iota, cap, hn, order, overflow);
    if (st != SKW_OK) return st;
    if (overflow) {
      if (attempt >= 3 || cap >= 0x40000000u) return fail(SKW_OUT_OF_MEMORY, "device tree build: too many nodes");
      cap *= 4;
      continue;
    }
    return layout(pt, n, hn, order, out);
  }
}

}  // namespace skw
