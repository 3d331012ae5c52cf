// Photon kd-tree built on the device (skw_photon_map_build, and the maps of a shot photon set):
// the host lists the tree's node ranges level by level (they depend on n only), the kernels of
// kd_build.hip arrange the items and write the node words, and the map's arrays are filled on the
// device; the host keeps copies as for a host-built map.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace {

// Height of the subtree of a range of s items (the tree's shape depends on s only).
uint32_t subtree_height_rec(uint64_t s) {
  if (s <= 1) return 0;
  const uint64_t left = s / 2, right = s - s / 2 - 1;
  uint32_t h = subtree_height_rec(left);
  if (right > 0) h = std::max(h, subtree_height_rec(right));
  return h + 1;
}

// Height of the subtree over s items (the plan asks it for every node of at most kKdSeqMax items).
uint32_t subtree_height(uint64_t s) {
  static const std::vector<uint32_t> table = [] {
    std::vector<uint32_t> t(kKdSeqMax + 1);
    for (uint64_t i = 0; i <= kKdSeqMax; ++i) t[i] = subtree_height_rec(i);
    return t;
  }();
  return s <= kKdSeqMax ? table[s] : subtree_height_rec(s);
}

struct Plan {
  std::vector<uint64_t> ranges;  // every list, concatenated: (begin, end) pairs
  struct Level {
    size_t big_off, big_n, mid_off, mid_n;
    uint64_t mid_max;
  };
  std::vector<Level> levels;
  size_t small_off = 0, small_n = 0;
  uint32_t max_depth = 0;
};

Plan plan_levels(uint64_t n) {
  Plan p;
  std::vector<uint64_t> cur, small;
  if (n > 0) cur = {0, n};
  std::vector<uint32_t> small_level;
  for (uint32_t level = 0; !cur.empty(); ++level) {
    std::vector<uint64_t> big, mid, next;
    next.reserve(2 * cur.size());
    uint64_t mmax = 0;
    for (size_t i = 0; i < cur.size(); i += 2) {
      const uint64_t b = cur[i], e = cur[i + 1], s = e - b;
      if (s <= kKdSeqMax) {
        small.push_back(b);
        small.push_back(e);
        p.max_depth = std::max<uint32_t>(p.max_depth, level + subtree_height(s));
        continue;
      }
      if (s > kKdLdsItems) {
        big.push_back(b);
        big.push_back(e);
      } else {
        mid.push_back(b);
        mid.push_back(e);
        mmax = std::max(mmax, s);
      }
      p.max_depth = std::max(p.max_depth, level);
      const uint64_t m = (b + e) / 2;
      if (b < m) {
        next.push_back(b + 1);
        next.push_back(m + 1);
      }
      if (m + 1 < e) {
        next.push_back(m + 1);
        next.push_back(e);
      }
    }
    Plan::Level L;
    L.big_off = p.ranges.size() / 2;
    L.big_n = big.size() / 2;
    p.ranges.insert(p.ranges.end(), big.begin(), big.end());
    L.mid_off = p.ranges.size() / 2;
    L.mid_n = mid.size() / 2;
    L.mid_max = mmax;
    p.ranges.insert(p.ranges.end(), mid.begin(), mid.end());
    if (L.big_n + L.mid_n > 0) p.levels.push_back(L);
    cur.swap(next);
  }
  p.small_off = p.ranges.size() / 2;
  p.small_n = small.size() / 2;
  p.ranges.insert(p.ranges.end(), small.begin(), small.end());
  return p;
}

// Levels with fewer large nodes than this run each introselect step with many workgroups per
// node (chunks of kKdChunk positions); the others with one workgroup per node.
constexpr size_t kMultiNodes = 64;

// The large nodes of one level, many workgroups per node: introselect steps until every node's
// subrange fits in LDS (one small read-back of the node states per step), then each node's tail.
hipError_t multi_level(const KdBuildArgs &a, const Plan &plan, const Plan::Level &L, hipStream_t st) {
  const uint32_t nn = (uint32_t)L.big_n;
  std::vector<KdStepNode> nodes(nn);
  std::vector<uint32_t> chunk_node;
  for (uint32_t i = 0; i < nn; ++i) {
    KdStepNode &nd = nodes[i];
    std::memset(&nd, 0, sizeof(nd));
    nd.b = (int64_t)plan.ranges[2 * (L.big_off + i)];
    nd.e = (int64_t)plan.ranges[2 * (L.big_off + i) + 1];
    nd.chunk0 = (uint32_t)chunk_node.size();
    nd.nchunks = (uint32_t)((uint64_t)(nd.e - nd.b + kKdChunk - 1) / kKdChunk);
    for (uint32_t c = 0; c < nd.nchunks; ++c) chunk_node.push_back(i);
  }
  const uint32_t nc = (uint32_t)chunk_node.size();
  KdStepNode *d_nodes = nullptr;
  uint32_t *d_chunk = nullptr;
  float *d_work = nullptr;
  hipError_t e = pool_alloc((void **)&d_nodes, nn * sizeof(KdStepNode), st);
  if (e == hipSuccess) e = pool_alloc((void **)&d_chunk, nc * sizeof(uint32_t), st);
  if (e == hipSuccess) e = pool_alloc((void **)&d_work, (size_t)nc * 8 * sizeof(float), st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_nodes, nodes.data(), nn * sizeof(KdStepNode), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_chunk, chunk_node.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_kd_step_init(a, d_nodes, nn, d_chunk, nc, d_work, st);
  // Each step shrinks an active subrange, and the depth budget bounds the steps. Steps are
  // enqueued kStepGroup at a time between read-backs: a node that has left the loop stays out and
  // every step kernel skips it, so the steps past the last active one change nothing.
  constexpr int kStepGroup = 8;
  for (int step = 0; e == hipSuccess && step < 256; step += kStepGroup) {
    for (int g = 0; g < kStepGroup && e == hipSuccess; ++g) {
      e = launch_kd_step_begin(a, d_nodes, nn, st);
      if (e == hipSuccess) e = launch_kd_step_rest(a, d_nodes, nn, d_chunk, nc, d_work, st);
    }
    if (e == hipSuccess) e = launch_kd_step_begin(a, d_nodes, nn, st);
    if (e == hipSuccess) e = hipMemcpyAsync(nodes.data(), d_nodes, nn * sizeof(KdStepNode), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) break;
    bool any = false;
    for (const KdStepNode &nd : nodes) any = any || nd.active;
    if (!any) break;
    e = launch_kd_step_rest(a, d_nodes, nn, d_chunk, nc, d_work, st);
  }
  if (e == hipSuccess) e = launch_kd_step_tail(a, d_nodes, nn, st);
  pool_free(d_nodes, st);
  pool_free(d_chunk, st);
  pool_free(d_work, st);
  return e;
}

template <class T>
hipError_t download(std::vector<T> &dst, const T *src, uint64_t n, hipStream_t st) {
  dst.resize(n);
  if (n == 0) return hipSuccess;
  return hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, st);
}

}  // namespace

namespace skw {

skw_status photon_map_build_device(uint64_t n, const float *d_pos, const float *d_weight, const uint32_t *d_codes,
                                   const uint16_t *d_dir, const uint16_t *d_normal, skw_photon_map **out,
                                   hipStream_t st) {
  *out = nullptr;
  if (n >= (uint64_t)kNoChild) return fail(SKW_INVALID_ARG, "too many photons for the 29-bit child index");
  if (n > 0 && !d_pos) return fail(SKW_INVALID_ARG, "positions must not be NULL");
  std::unique_ptr<skw_photon_map> m(new skw_photon_map());
  SKW_HIP(hipGetDevice(&m->device), "hipGetDevice");
  const Plan plan = plan_levels(n);
  KdBuildArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  uint64_t *d_ranges = nullptr;
  char *work = nullptr;
  const size_t items_bytes = ((n * 16 + 255) / 256) * 256, words_bytes = ((n * 4 + 255) / 256) * 256;
  hipError_t e = hipSuccess;
  if (n > 0) {
    e = pool_alloc((void **)&work, items_bytes + 3 * words_bytes, st);
    if (e == hipSuccess) e = pool_alloc((void **)&d_ranges, std::max<size_t>(plan.ranges.size(), 1) * 8, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_ranges, plan.ranges.data(), plan.ranges.size() * 8, hipMemcpyHostToDevice, st);
    a.items = work;
    a.word = reinterpret_cast<uint32_t *>(work + items_bytes);
    a.scr_l = reinterpret_cast<uint32_t *>(work + items_bytes + words_bytes);
    a.scr_r = reinterpret_cast<uint32_t *>(work + items_bytes + 2 * words_bytes);
    m->pooled = true;
    if (e == hipSuccess) e = pool_alloc((void **)&m->d_nodes, n * sizeof(DevKdNode), st);
    if (e == hipSuccess) e = pool_alloc((void **)&m->d_photons, n * sizeof(DevPhoton), st);
    if (e == hipSuccess) e = pool_alloc((void **)&m->d_perm, n * sizeof(uint32_t), st);
    if (e == hipSuccess) e = pool_alloc((void **)&m->d_psplit, n * sizeof(float), st);
    if (e == hipSuccess) e = launch_kd_init(a, d_pos, st);
    for (const Plan::Level &L : plan.levels) {
      if (e != hipSuccess) break;
      if (L.big_n > 0 && L.big_n < kMultiNodes) e = multi_level(a, plan, L, st);
      else e = launch_kd_big(a, d_ranges + 2 * L.big_off, (uint32_t)L.big_n, st);
      if (e == hipSuccess) e = launch_kd_mid(a, d_ranges + 2 * L.mid_off, (uint32_t)L.mid_n, L.mid_max, st);
    }
    if (e == hipSuccess) e = launch_kd_small(a, d_ranges + 2 * plan.small_off, (uint32_t)plan.small_n, st);
    if (e == hipSuccess)
      e = launch_kd_finish(a, d_weight, d_codes, d_dir, d_normal, m->d_nodes, m->d_photons, m->d_perm, m->d_psplit,
                           st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    pool_free(work, st);
    pool_free(d_ranges, st);
  }
  if (e != hipSuccess) return hip_fail(e, "device kd build");
  m->n = n;
  m->max_depth = plan.max_depth;
  m->host.max_depth = plan.max_depth;
  m->host_ready = n == 0;
  m->pooled = true;
  SKW_HIP(shared_cdir(&m->d_cdir), "direction table");
  m->cdir_shared = true;
  m->on_device = true;
  *out = m.release();
  return SKW_OK;
}

hipError_t shared_cdir(double **out) {
  static std::mutex mu;
  static double *tables[64] = {nullptr};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!tables[dev]) {
    std::vector<double> table(cdir_table(), cdir_table() + 3 * 65536);
    double *d = nullptr;
    e = upload(&d, table);
    if (e != hipSuccess) {
      if (d) (void)hipFree(d);
      return e;
    }
    tables[dev] = d;  // kept for the life of the process
  }
  *out = tables[dev];
  return hipSuccess;
}

hipError_t side_stream(hipStream_t *out) {
  static std::mutex mu;
  static hipStream_t streams[64] = {nullptr};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev]) {
    e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
    if (e != hipSuccess) return e;
  }
  *out = streams[dev];
  return hipSuccess;
}

const KdLayout &photon_map_host(const skw_photon_map *map) {
  skw_photon_map *m = const_cast<skw_photon_map *>(map);
  std::lock_guard<std::mutex> lock(m->host_mu);
  if (!m->host_ready) {
    KdLayout &h = m->host;
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(m->device);
    if (e == hipSuccess) e = download(h.nodes, m->d_nodes, m->n, nullptr);
    if (e == hipSuccess) e = download(h.photons, m->d_photons, m->n, nullptr);
    if (e == hipSuccess) e = download(h.perm, m->d_perm, m->n, nullptr);
    if (e == hipSuccess) e = download(h.psplit, m->d_psplit, m->n, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    (void)hipSetDevice(prev);
    if (e == hipSuccess) m->host_ready = true;
    else h.nodes.clear();  // callers see the size mismatch
  }
  return m->host;
}

}  // namespace skw

extern "C" {

skw_status skw_photon_map_build(uint64_t n, const float *d_pos, const float *d_weight, const uint16_t *d_dir,
                                const uint16_t *d_normal, skw_photon_map **out, void *stream) {
  if (!out) return fail(SKW_INVALID_ARG, "out must not be NULL");
  *out = nullptr;
  if (device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  return photon_map_build_device(n, d_pos, d_weight, nullptr, d_dir, d_normal, out,
                                 reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
