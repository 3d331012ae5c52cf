// Triangle tree built on the device (skw_scene_create with SKW_SCENE_DEVICE_BUILD): the host
// prepares the world-space triangles and their boxes exactly as the host builder does, the
// kernels of tri_build.hip build the tree level by level (node boxes, split search, the in-place
// three-way partition), and the host lays the nodes out depth first as the host builder's
// place() does, so the scene is node for node the one the host builder makes
// (TriangleAccelerator.cpp:22-147, 505-729).
#include <hip/hip_runtime.h>

#include <algorithm>
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

// Lays out node `id` at index idx, its children as one block appended at the end (place() of the
// host builder).
void place_dev(const std::vector<TriBuildNode> &hn, uint32_t id, uint32_t idx, TreeLayout &out) {
  const TriBuildNode &n = hn[id];
  {
    DevNode &d = out.nodes[idx];
    for (int i = 0; i < 3; ++i) {
      d.bmin[i] = n.bmin[i];
      d.bmax[i] = n.bmax[i];
    }
    if (n.first_child == kNoChild) {
      d.a = n.tb;
      d.b = (n.te - n.tb) << 3;
      out.leaves++;
      return;
    }
  }
  const uint32_t flags = n.info & 3u, mask = (n.info >> 2) & 7u, nch = (uint32_t)__builtin_popcount(mask);
  const uint32_t first = (uint32_t)out.nodes.size();
  out.nodes[idx].a = first;
  out.nodes[idx].b = flags | (mask << 2);
  out.nodes.resize(out.nodes.size() + nch);
  for (uint32_t k = 0; k < nch; ++k) place_dev(hn, n.first_child + k, first + k, out);
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
  for (float v : pt.box)
    if (!std::isfinite(v)) return fail(SKW_UNSUPPORTED, "the device tree build needs finite vertices");
  const uint32_t n = (uint32_t)n64;
  const uint32_t small_max = (uint32_t)std::min(256, std::max(1, env_int("SKW_TRIBUILD_SMALL_MAX", 32)));

  Buffers B;
  SKW_HIP(hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking), "hipStreamCreate");
  TriBuildArgs a;
  std::memset(&a, 0, sizeof(a));
  SKW_HIP(B.alloc(&a.box, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&a.box2, 6 * (size_t)n), "tree build allocation");
  SKW_HIP(B.alloc(&a.ref, n), "tree build allocation");
  SKW_HIP(B.alloc(&a.ref2, n), "tree build allocation");
  uint32_t **scratch[9] = {&a.s_cls, &a.s_rf, &a.s_nb, &a.s_G, &a.s_Rpos, &a.s_order, &a.s_elem, &a.s_ptr, &a.s_dest};
  for (uint32_t **s : scratch) SKW_HIP(B.alloc(s, n), "tree build allocation");
  uint32_t *d_counts = nullptr;  // node count, next big, next small
  SKW_HIP(B.alloc(&d_counts, 4), "tree build allocation");
  std::vector<uint32_t> iota(n);
  for (uint32_t i = 0; i < n; ++i) iota[i] = i;
  SKW_HIP(hipMemcpyAsync(a.box, pt.box.data(), pt.box.size() * sizeof(float), hipMemcpyHostToDevice, B.st),
          "upload boxes");
  SKW_HIP(hipMemcpyAsync(a.ref, iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, B.st), "upload ids");

  // nodes and the four lists (current / next, big / small), grown before a level could overflow
  uint32_t cap = std::max<uint32_t>(1024, n / 2 + 64);
  TriBuildNode *d_nodes = nullptr;
  uint32_t *lists[4] = {nullptr, nullptr, nullptr, nullptr};
  SKW_HIP(B.alloc(&d_nodes, cap), "tree build allocation");
  for (uint32_t *&l : lists) SKW_HIP(B.alloc(&l, cap), "tree build allocation");
  TriBuildNode root;
  std::memset(&root, 0, sizeof(root));
  root.tb = 0;
  root.te = n;
  root.first_child = kNoChild;
  root.info = 0;
  uint32_t h_counts[4] = {1, 0, 0, 0};
  const uint32_t zero = 0;
  SKW_HIP(hipMemcpyAsync(d_nodes, &root, sizeof(root), hipMemcpyHostToDevice, B.st), "upload root");
  SKW_HIP(hipMemcpyAsync(n > small_max ? lists[0] : lists[1], &zero, sizeof(zero), hipMemcpyHostToDevice, B.st),
          "upload root list");
  uint32_t count = 1, nbig = n > small_max ? 1u : 0u, nsmall = n > small_max ? 0u : 1u;
  while (nbig + nsmall > 0) {
    const uint64_t need = (uint64_t)count + 3ull * (nbig + nsmall);
    if (need > 0xFFFFFFF0ull) return fail(SKW_OUT_OF_MEMORY, "device tree build: too many nodes");
    if (need > cap) {
      const uint32_t ncap = (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, std::max<uint64_t>(2ull * cap, need));
      TriBuildNode *nn = nullptr;
      SKW_HIP(B.alloc(&nn, ncap), "tree build allocation");
      SKW_HIP(hipMemcpyAsync(nn, d_nodes, (size_t)count * sizeof(TriBuildNode), hipMemcpyDeviceToDevice, B.st),
              "grow nodes");
      B.release(d_nodes);
      d_nodes = nn;
      for (int l = 0; l < 4; ++l) {
        uint32_t *nl = nullptr;
        SKW_HIP(B.alloc(&nl, ncap), "tree build allocation");
        const uint32_t keep = l == 0 ? nbig : (l == 1 ? nsmall : 0u);
        if (keep)
          SKW_HIP(hipMemcpyAsync(nl, lists[l], (size_t)keep * sizeof(uint32_t), hipMemcpyDeviceToDevice, B.st),
                  "grow lists");
        B.release(lists[l]);
        lists[l] = nl;
      }
      cap = ncap;
    }
    h_counts[0] = count;
    h_counts[1] = h_counts[2] = 0;
    SKW_HIP(hipMemcpyAsync(d_counts, h_counts, 3 * sizeof(uint32_t), hipMemcpyHostToDevice, B.st), "set counts");
    a.nodes = d_nodes;
    a.node_count = d_counts;
    a.next_big = lists[2];
    a.next_big_n = d_counts + 1;
    a.next_small = lists[3];
    a.next_small_n = d_counts + 2;
    a.small_max = small_max;
    a.list = lists[0];
    a.list_n = nbig;
    SKW_HIP(launch_tri_build_level(a, B.st), "tree build level");
    a.list = lists[1];
    a.list_n = nsmall;
    SKW_HIP(launch_tri_build_small(a, B.st), "tree build small nodes");
    SKW_HIP(hipMemcpyAsync(h_counts, d_counts, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, B.st), "read counts");
    SKW_HIP(hipStreamSynchronize(B.st), "tree build level");
    if (h_counts[0] > count + 3ull * (nbig + nsmall))
      return fail(SKW_HIP_ERROR, "device tree build: node count out of range");
    count = h_counts[0];
    nbig = h_counts[1];
    nsmall = h_counts[2];
    std::swap(lists[0], lists[2]);
    std::swap(lists[1], lists[3]);
  }
  std::vector<TriBuildNode> hn(count);
  std::vector<uint32_t> order(n);
  SKW_HIP(hipMemcpyAsync(hn.data(), d_nodes, (size_t)count * sizeof(TriBuildNode), hipMemcpyDeviceToHost, B.st),
          "download nodes");
  SKW_HIP(hipMemcpyAsync(order.data(), a.ref, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, B.st),
          "download triangle order");
  SKW_HIP(hipStreamSynchronize(B.st), "download tree");

  out.nodes.assign(1, DevNode());
  out.leaves = 0;
  place_dev(hn, 0, 0, out);
  out.root = 0;
  uint32_t depth = 0;
  for (uint32_t i = 0; i < count; ++i) depth = std::max(depth, hn[i].info >> 16);
  out.max_depth = depth;
  out.stack_need = depth + 1;
  for (int i = 0; i < 3; ++i) {
    out.bounds[i] = hn[0].bmin[i];
    out.bounds[3 + i] = hn[0].bmax[i];
  }
  out.tris.resize(n);
  for (uint32_t pos = 0; pos < n; ++pos) {
    const uint32_t id = order[pos];
    if (id >= n) return fail(SKW_HIP_ERROR, "device tree build: triangle order out of range");
    DevTri &t = out.tris[pos];
    std::memcpy(t.v0, &pt.tv[9 * (size_t)id], 3 * sizeof(float));
    std::memcpy(t.v1, &pt.tv[9 * (size_t)id + 3], 3 * sizeof(float));
    std::memcpy(t.v2, &pt.tv[9 * (size_t)id + 6], 3 * sizeof(float));
    t.prim = pt.tprim[id];
    t.tri = pt.ttri[id];
    t.pad = 0;
  }
  return SKW_OK;
}

}  // namespace skw
